"""Input decks (iceicle_amd.deck): the mapping of a deck's tables onto the
engine's mesh, discretization and solver descriptors (CPU), and decks run
on the device against the published 1D heat-equation study and against the
same problems set up through the analytic-field entry points (GPU).  The
decks here are written for these tests in the reference's deck format."""
import math

import numpy as np
import pytest

from iceicle_amd import abi, deck

HEAT_DECK = """
-- heat equation through Burgers (mu = 1, no advection), Dirichlet 0 at both
-- ends of [0, 2 pi], initial state sin x, RK3-TVD at Fourier number 1e-4
local nelem = {nelem}
local mu = 1.0
local tfinal = {tfinal}
return {{
  ndim = 1,
  uniform_mesh = {{
    nelem = {{ nelem }},
    bounding_box = {{ min = {{ 0.0 }}, max = {{ 2 * math.pi }} }},
    boundary_conditions = {{ types = {{ "dirichlet", "dirichlet" }}, flags = {{ 0, 0 }} }},
    geometry_order = 1,
  }},
  fespace = {{ basis = "legendre", quadrature = "gauss", order = {order} }},
  conservation_law = {{ name = "burgers", mu = mu }},
  initial_condition = function(x) return math.sin(x) end,
  boundary_conditions = {{ dirichlet = {{ 0.0 }} }},
  solver = {{ type = "rk3-tvd", dt = 1e-4 * (2 * math.pi / nelem) ^ 2, tfinal = tfinal, ivis = 100000 }},
  post = {{
    exact_solution = function(x) return math.sin(x) * math.exp(-mu * tfinal) end,
    tasks = {{ "l2_error", "l1_error", "linf_error" }},
  }},
}}
"""

MMS_2D_DECK = """
local mu = 0.05
local function exact(x, y) return math.sin(math.pi * x) * math.cos(0.5 * math.pi * y) + 0.25 * x * y end
return {
  ndim = 2,
  uniform_mesh = {
    nelem = { 6, 5 },
    bounding_box = { min = { 0.0, 0.0 }, max = { 1.0, 1.0 } },
    boundary_conditions = {
      types = { "dirichlet", "neumann", "dirichlet", "dirichlet" },
      flags = { 0, 0, 1, 0 },
    },
  },
  fespace = { order = 3 },
  conservation_law = {
    name = "burgers", mu = mu, a_adv = { 0.3, -0.2 }, b_adv = { 0.0, 0.0 },
    source = function(x, y) return 1.0 + x * y end,
  },
  initial_condition = exact,
  boundary_conditions = {
    dirichlet = { exact, 0.5 },
    neumann = { function(x, y) return x - y end },
  },
}
"""


def test_uniform_mesh_and_params_mapping():
    d = deck.Deck(source=MMS_2D_DECK, out=lambda s: None)
    m = d.mesh()
    assert m.ndim == 2 and list(m.nelem)[:2] == [6, 5]
    assert [m.bctype[k] for k in range(4)] == [abi.BC["DIRICHLET"], abi.BC["NEUMANN"], abi.BC["DIRICHLET"],
                                               abi.BC["DIRICHLET"]]
    assert [m.bcflag[k] for k in range(4)] == [0, 0, 1, 0]
    p = d.params()
    assert p.flux == abi.FLUX_BURGERS and p.mu == 0.05
    assert list(p.a)[:2] == [0.3, -0.2] and list(p.b)[:2] == [0.0, 0.0]


def test_directional_nodes_and_user_mesh_chain():
    src = """
    return { ndim = 2, uniform_mesh = {
      directional_nodes = { { 0.0, 0.1, 0.4, 1.0 }, { -1.0, 2.0 } },
      boundary_conditions = { types = { "dirichlet", "riemann", "slip wall", "wall" }, flags = { 0, 0, 0, 1 } } },
      fespace = { order = 1 }, conservation_law = { name = "euler" } }
    """
    d = deck.Deck(source=src, out=lambda s: None)
    m = d.mesh()
    assert list(m.nelem)[:2] == [3, 1]
    c = m._coord_owner
    assert c.shape == (8, 2) and np.array_equal(c[:, 0], np.repeat([0.0, 0.1, 0.4, 1.0], 2))
    assert [m.bctype[k] for k in range(4)] == [abi.BC["DIRICHLET"], abi.BC["RIEMANN"], abi.BC["SLIP_WALL"],
                                               abi.BC["NO_SLIP_ISOTHERMAL"]]
    one = """
    local coord = {}
    for i = 0, 4 do coord[#coord + 1] = { (i / 4) ^ 2 } end
    local elements = {}
    for i = 0, 3 do elements[#elements + 1] = { domain_type = "hypercube", order = 1, nodes = { i, i + 1 } } end
    return { ndim = 1, user_mesh = { coord = coord, elements = elements, boundary_faces = {
        { bc_type = "dirichlet", bc_flag = 0, nodes = { 0 } }, { bc_type = "extrapolation", bc_flag = 0, nodes = { 4 } } } },
      fespace = { order = 2 }, conservation_law = { name = "burgers", mu = 0.1 } }
    """
    m1 = deck.Deck(source=one, out=lambda s: None).mesh()
    assert m1.nelem[0] == 4
    assert np.allclose(m1._coord_owner[:, 0], (np.arange(5) / 4) ** 2)
    assert [m1.bctype[0], m1.bctype[1]] == [abi.BC["DIRICHLET"], abi.BC["EXTRAPOLATION"]]


def test_free_stream_conversion():
    src = """
    local gamma = 1.4
    local rho, u, v, p = 1.2, 3.0, 4.0, 2.0
    return { ndim = 2, uniform_mesh = { nelem = { 2, 2 }, bounding_box = { min = { 0, 0 }, max = { 1, 1 } },
      boundary_conditions = { types = { "riemann", "riemann", "riemann", "riemann" }, flags = { 0, 0, 0, 0 } } },
      fespace = { order = 1 },
      conservation_law = { name = "euler", free_stream = { rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * 25 } } }
    """
    p = deck.Deck(source=src, out=lambda s: None).params()
    assert p.flux == abi.FLUX_EULER
    assert abs(p.freestream[0] - 1.2) < 1e-15 and abs(p.freestream[1] - 5.0) < 1e-14
    assert abs(p.freestream[2] - 0.6) < 1e-15 and abs(p.freestream[3] - 0.8) < 1e-15
    assert abs(p.freestream[5] * 1.2 * p.Rgas - 2.0) < 1e-12   # p = rho Rgas T


@pytest.mark.parametrize("entry,msg", [
    ("mdg = {}", "'mdg'"),
    ("gmsh = { file = 'x.msh' }", "'gmsh'"),
])
def test_unsupported_entries_are_refused(entry, msg):
    src = "return { ndim = 1, " + entry + " }"
    with pytest.raises(deck.DeckError, match=msg):
        deck.Deck(source=src, out=lambda s: None)


def test_unknown_boundary_name_and_law():
    src = """return { ndim = 1, uniform_mesh = { nelem = { 2 }, bounding_box = { min = { 0 }, max = { 1 } },
      boundary_conditions = { types = { "dirichlet", "outflow" }, flags = { 0, 0 } } },
      conservation_law = { name = "spacetime-burgers" } }"""
    d = deck.Deck(source=src, out=lambda s: None)
    with pytest.raises(deck.DeckError, match="outflow"):
        d.mesh()
    with pytest.raises(deck.DeckError, match="spacetime-burgers"):
        d.params()


# ---------------------------------------------------------------- on the GPU

PUBLISHED_DDG_P2_NX8 = 0.00190644295   # doc/sine_diffusion_1d_study.md:43 (standard DDG, nx 8, P2)


@pytest.mark.gpu
def test_heat_deck_reproduces_the_published_study():
    """The deck form of the published study's case (sine_diffusion_1d, nx 8,
    P2, to t = 1) gives its L2 error; the Legendre basis of the study and the
    Lagrange basis here span the same space."""
    pytest.importorskip("torch")
    lines = []
    r = deck.Deck(source=HEAT_DECK.format(nelem=8, order=2, tfinal=1.0), out=lines.append).run()
    err = r["post"]["l2_error"]
    print(f"deck heat nx 8 P2: L2 {err:.10g} (published {PUBLISHED_DDG_P2_NX8}), steps {r['solver']['steps']}",
          flush=True)
    assert abs(r["solver"]["time"] - 1.0) < 1e-12
    assert abs(err - PUBLISHED_DDG_P2_NX8) / PUBLISHED_DDG_P2_NX8 < 1e-6
    assert any(l.startswith("L2 error:") for l in lines)
    assert r["post"]["l1_error"] > 0 and len(r["post"]["linf_error"]) == 1


@pytest.mark.gpu
def test_deck_callbacks_equal_field_setup():
    """A deck's Lua initial condition, Dirichlet / Neumann callbacks and
    source give the residual of the same problem set up with the values
    computed here (projection, boundary table, source load) on the device,
    and the boundary table matches an analytic constant Dirichlet field."""
    torch = pytest.importorskip("torch")
    from iceicle_amd import engine
    d = deck.Deck(source=MMS_2D_DECK, out=lambda s: None)
    d.mesh_desc = d.mesh()
    space = engine.FESpace(d.mesh_desc, 3)
    disc = engine.ConservationLawDDG(space, d.params())
    d.boundary_and_source(space, disc)
    u = torch.zeros(disc.size, dtype=torch.float64, device="cuda")
    d.initial_condition(space, disc, u)
    r = torch.empty_like(u)
    engine.form_residual(space, disc, u, r)
    # the same by hand
    xq = space.qp_coords().reshape(-1, 2)

    def exact(x, y):
        return np.sin(np.pi * x) * np.cos(0.5 * np.pi * y) + 0.25 * x * y

    p = d.params()
    disc2 = engine.ConservationLawDDG(space, p)
    xb, bt, bf = space.bface_coords()
    bv = np.zeros((xb.shape[0], space.nfqp, 1))
    for b in range(xb.shape[0]):
        if bt[b] == abi.BC["DIRICHLET"]:
            bv[b, :, 0] = exact(xb[b, :, 0], xb[b, :, 1]) if bf[b] == 0 else 0.5
        elif bt[b] == abi.BC["NEUMANN"]:
            bv[b, :, 0] = xb[b, :, 0] - xb[b, :, 1]
    disc2.set_boundary_values(torch.from_numpy(bv.ravel()).cuda())
    disc2.set_source_values(torch.from_numpy((1.0 + xq[:, 0] * xq[:, 1]).copy()).cuda())
    u2 = torch.zeros_like(u)
    engine.project_values(disc2, torch.from_numpy(exact(xq[:, 0], xq[:, 1]).copy()).cuda(), u2)
    r2 = torch.empty_like(u2)
    engine.form_residual(space, disc2, u2, r2)
    scale = float(r2.abs().max())
    assert scale > 0
    assert float((u - u2).abs().max()) <= 1e-14 * float(u2.abs().max())
    assert float((r - r2).abs().max()) <= 1e-13 * scale
    # projection of values against the analytic-field projection
    fld = abi.Field.make("SINE", [1.0, 1.0, 0.0]) if hasattr(abi.Field, "make") else None
    if fld is not None:
        m1 = abi.mesh_desc(1, [5], [0.0], [2 * math.pi], ["DIRICHLET", "DIRICHLET"])
        s1 = engine.FESpace(m1, 3)
        d1 = engine.ConservationLawDDG(s1, abi.default_params(abi.FLUX_BURGERS))
        ua = torch.zeros(d1.size, dtype=torch.float64, device="cuda")
        engine.project(d1, fld, ua)
        xs = s1.qp_coords().reshape(-1)
        ub = torch.zeros_like(ua)
        engine.project_values(d1, torch.from_numpy(np.sin(xs)).cuda(), ub)
        assert float((ua - ub).abs().max()) <= 1e-14
