"""Run the reference's Lua input decks on the device engine (the driver of
conservation_law.cpp:80-185 with solvers_lua_interface.hpp:402-694,
mesh_lua_interface.hpp:23-97, initialization.hpp:37-104 and
bc_lua_interface.hpp:71-160).

A deck is executed by the Lua-subset interpreter (`iceicle_amd.lua`); its
returned table is mapped onto the engine:

* mesh: `uniform_mesh` with `bounding_box` and `nelem`, or with
  `directional_nodes`; boundary `types` by the reference's names
  (get_bc_from_name, face.hpp:126-166: "dirichlet", "neumann",
  "extrapolation", "spacetime-future", "spacetime-past", "slip wall",
  "isothermal" / "no-slip isothermal" / "wall" / "general wall" (all no-slip
  isothermal), "riemann" / "characteristic"; a name the reference does not
  know becomes an interior face there and is refused here) and `flags`;
  `geometry_order` 1.  A 1D `user_mesh` whose elements form one chain is
  taken as `directional_nodes`.
* `fespace.order` 1-4 (the device kernels' range); `basis` "lagrange" or
  "legendre" (the same DG space: the coefficients here are Lagrange ones);
  `quadrature` "gauss".
* `conservation_law`: "burgers" (`mu`, `a_adv`, `b_adv`, `source`),
  "euler" and "navier-stokes" (the reference's driver has those branches
  commented out, conservation_law.cpp:287-345; here they run with the
  defaults of ice_disc_params_default and the deck's `free_stream`, `mu`,
  `gamma`, `Pr` when given); `sigma_ic`, `interior_penalty`.
* `initial_condition`: a function of the coordinates (a number for one
  equation, a table otherwise) or "zero", L2-projected on the device from
  its values at the volume quadrature points; or `restart` (read_restart).
* `boundary_conditions.dirichlet` / `.neumann`: per flag a number (every
  equation) or a function of the coordinates, evaluated at the boundary face
  quadrature points.
* `solver`: "explicit_euler", "rk3-ssp", "rk3-tvd" with `dt` or `cfl` and
  `tfinal` or `ntime`, `ivis`; "newton" with `tau_abs`, `tau_rel`, `kmax`,
  `linesearch` (`type` "none", "wolfe" / "cubic", "corrigan").
* `output.writer` "dat" (1D) or "vtu" (2D/3D): the initial condition and
  every `ivis` steps.
* `post.tasks` with `post.exact_solution`: "l2_error", "l1_error",
  "linf_error" (l2_error.hpp:37-250), "plot_exact_projection", and
  "ic_residual" (the interface error, ice_interface_error), which needs no
  exact solution.
Unsupported entries (MDG geometry cycles, gmsh / mixed / user meshes in 2D
and 3D, spacetime Burgers, the Gauss-Newton and matrix-free solvers, mesh
perturbation and management) raise DeckError naming the entry.
"""
import os

import numpy as np

from . import abi, engine, io, lua


class DeckError(RuntimeError):
    pass


BC_NAMES = {
    "dirichlet": "DIRICHLET", "neumann": "NEUMANN", "extrapolation": "EXTRAPOLATION",
    "spacetime-future": "SPACETIME_FUTURE", "spacetime-past": "SPACETIME_PAST", "slip wall": "SLIP_WALL",
    "isothermal": "NO_SLIP_ISOTHERMAL", "no-slip isothermal": "NO_SLIP_ISOTHERMAL",
    "wall": "NO_SLIP_ISOTHERMAL", "general wall": "NO_SLIP_ISOTHERMAL", "riemann": "RIEMANN",
    "characteristic": "RIEMANN", "periodic": "PERIODIC",
}


def _get(t, *keys, default=None):
    for k in keys:
        if not isinstance(t, lua.LuaTable):
            return default
        t = t.get(k)
        if t is None:
            return default
    return t


def _array(t, what):
    if not isinstance(t, lua.LuaTable):
        raise DeckError(f"{what} must be a table")
    return t.array()


class Deck:
    """A deck loaded from a file or a source string; `run()` executes it."""

    def __init__(self, path=None, source=None, out=print, workdir=None):
        if (path is None) == (source is None):
            raise DeckError("give a deck path or its source")
        self.out = out
        self.path = path
        self.workdir = workdir or (os.path.dirname(os.path.abspath(path)) if path else os.getcwd())
        src = open(path).read() if path else source
        r = lua.Interpreter(out=out).run(src)
        if not r or not isinstance(r[0], lua.LuaTable):
            raise DeckError("the deck must return a configuration table")
        self.cfg = r[0]
        for key in ("mdg", "gmsh", "mixed_uniform_mesh", "burgers_mesh", "mesh_perturbation", "mesh_management"):
            if self.cfg.get(key) is not None:
                raise DeckError(f"'{key}' is not supported by this driver")
        self.ndim = int(self.cfg.get("ndim") or 0)
        if self.ndim not in (1, 2, 3):
            raise DeckError("ndim must be 1, 2 or 3")

    # ---- mesh
    def mesh(self):
        um = self.cfg.get("uniform_mesh")
        usr = self.cfg.get("user_mesh")
        D = self.ndim
        if um is None and usr is not None:
            return self._user_mesh_1d(usr)
        if um is None:
            raise DeckError("the deck has no uniform_mesh")
        if int(_get(um, "geometry_order", default=1)) != 1:
            raise DeckError("geometry_order above 1 is not supported")
        types = [str(x).strip().lower() for x in _array(_get(um, "boundary_conditions", "types"), "types")]
        flags = [int(x) for x in _array(_get(um, "boundary_conditions", "flags"), "flags")]
        if len(types) != 2 * D or len(flags) != 2 * D:
            raise DeckError(f"boundary_conditions needs {2 * D} types and flags")
        bct = []
        for t in types:
            if t not in BC_NAMES:
                raise DeckError(f"unknown boundary condition name {t!r}")
            bct.append(BC_NAMES[t])
        bb = um.get("bounding_box")
        if bb is not None:
            nel = [int(x) for x in _array(um.get("nelem"), "nelem")]
            lo = [float(x) for x in _array(bb.get("min"), "bounding_box.min")]
            hi = [float(x) for x in _array(bb.get("max"), "bounding_box.max")]
            return abi.mesh_desc(D, nel, lo, hi, bct, flags)
        dn = um.get("directional_nodes")
        if dn is None:
            raise DeckError("uniform_mesh needs bounding_box or directional_nodes")
        nodes = [np.array([float(v) for v in _array(t, "directional_nodes")]) for t in _array(dn, "directional_nodes")]
        return self._tensor_mesh(nodes, bct, flags)

    def _tensor_mesh(self, nodes, bct, flags):
        D = self.ndim
        if len(nodes) != D:
            raise DeckError("directional_nodes needs one node list per dimension")
        nel = [len(n) - 1 for n in nodes]
        grids = np.meshgrid(*nodes, indexing="ij")  # dimension 0 slowest
        coord = np.stack([g.ravel() for g in grids], axis=1)
        return abi.mesh_desc(D, nel, [n[0] for n in nodes], [n[-1] for n in nodes], bct, flags, coord=coord)

    def _user_mesh_1d(self, usr):
        if self.ndim != 1:
            raise DeckError("user_mesh is supported in 1D (one chain of elements) only")
        coord = [float(_array(c, "coord")[0]) if isinstance(c, lua.LuaTable) else float(c)
                 for c in _array(usr.get("coord"), "coord")]
        elems = _array(usr.get("elements"), "elements")
        chain = []
        for el in elems:
            nodes = _array(_get(el, "nodes"), "element nodes")
            chain.append((int(nodes[0]), int(nodes[1])))
        order = [chain[0][0]] + [b for _, b in chain]
        if any(chain[i][1] != chain[i + 1][0] for i in range(len(chain) - 1)):
            raise DeckError("the 1D user_mesh elements must form one chain in order")
        x = np.array([coord[i] for i in order])
        bct, flags = ["DIRICHLET", "DIRICHLET"], [0, 0]
        for bf in _array(usr.get("boundary_faces") or lua.LuaTable(), "boundary_faces"):
            name = str(_get(bf, "bc_type", default="dirichlet")).lower()
            node = int(_array(_get(bf, "nodes"), "boundary face nodes")[0])
            side = 0 if node == order[0] else 1 if node == order[-1] else None
            if side is None:
                raise DeckError("a 1D boundary face must sit at an end of the chain")
            if name not in BC_NAMES:
                raise DeckError(f"unknown boundary condition name {name!r}")
            bct[side] = BC_NAMES[name]
            flags[side] = int(_get(bf, "bc_flag", default=0))
        return self._tensor_mesh([x], bct, flags)

    # ---- discretization
    def params(self):
        cl = self.cfg.get("conservation_law")
        if cl is None:
            raise DeckError("the deck has no conservation_law table")
        name = str(cl.get("name")).lower()
        D = self.ndim
        if name == "burgers":
            p = abi.default_params(abi.FLUX_BURGERS)
            mu = cl.get("mu")
            if mu is not None:
                p.mu = float(mu)
            a = cl.get("a_adv")
            b = cl.get("b_adv")
            for d in range(3):
                p.a[d] = float(_array(a, "a_adv")[d]) if (a is not None and d < D) else 0.0
                p.b[d] = float(_array(b, "b_adv")[d]) if (b is not None and d < D) else 0.0
        elif name in ("euler", "navier-stokes"):
            p = abi.default_params(abi.FLUX_EULER if name == "euler" else abi.FLUX_NAVIER_STOKES)
            for key in ("mu", "gamma", "Pr"):
                v = cl.get(key)
                if v is not None:
                    setattr(p, key, float(v))
            fs = cl.get("free_stream")
            if fs is not None:
                # conservative (rho, rho v, rho E) -> (rho, |v|, v / |v|, T)
                # with p = (gamma - 1)(rho E - rho |v|^2 / 2) = rho Rgas T
                # (the unit reference quantities of ice_disc_params_default)
                vals = [float(v) for v in _array(fs, "free_stream")]
                if len(vals) != D + 2:
                    raise DeckError(f"free_stream needs {D + 2} conservative values")
                rho, mom, rhoE = vals[0], np.array(vals[1:1 + D]), vals[D + 1]
                vel = mom / rho
                vmag = float(np.linalg.norm(vel))
                pres = (p.gamma - 1.0) * (rhoE - 0.5 * rho * vmag ** 2)
                p.freestream[0] = rho
                p.freestream[1] = vmag
                for d in range(3):
                    p.freestream[2 + d] = (vel[d] / vmag if vmag > 0 else float(d == 0)) if d < D else 0.0
                p.freestream[5] = pres / (rho * p.Rgas)
        else:
            raise DeckError(f"conservation law {name!r} is not supported")
        for key in ("sigma_ic", "interior_penalty"):
            v = cl.get(key)
            if v is not None:
                setattr(p, key, float(v) if key == "sigma_ic" else int(bool(v)))
        return p

    # ---- callbacks as point values
    def _values(self, f, xs, neq, what):
        """[npts][neq] from a deck entry f: a number (every equation) or a
        function of the coordinates."""
        n = xs.shape[0]
        if isinstance(f, (int, float)) and not isinstance(f, bool):
            return np.full((n, neq), float(f))
        if f is None or not callable(f):
            raise DeckError(f"{what} must be a number or a function")
        return lua.vectorize(f, neq, *[xs[:, d] for d in range(self.ndim)])

    def initial_condition(self, space, disc, u):
        import torch
        rs = self.cfg.get("restart")
        if rs is not None:
            geo = io.HostGeometry(self.mesh_desc, space.order)
            coord, vals = io.read_restart(str(rs), self.ndim, geo.coord.shape[0], disc.size,
                                          directory=self.workdir)
            u.copy_(torch.from_numpy(vals))
            return
        ic = self.cfg.get("initial_condition")
        if ic is None or (isinstance(ic, str) and ic.lower() == "zero"):
            u.zero_()
            return
        if isinstance(ic, str):
            raise DeckError(f"initial condition {ic!r} is not known")
        xq = space.qp_coords().reshape(-1, self.ndim)
        vq = self._values(ic, xq, disc.neq, "initial_condition")
        engine.project_values(disc, torch.from_numpy(vq.ravel()).to(u.device), u)

    def boundary_and_source(self, space, disc):
        import torch
        bc = self.cfg.get("boundary_conditions")
        xb, bt, bf = space.bface_coords()
        if bc is not None and xb.shape[0] > 0:
            bv = np.zeros((xb.shape[0], space.nfqp, disc.neq))
            for kind, code in (("dirichlet", abi.BC["DIRICHLET"]), ("neumann", abi.BC["NEUMANN"])):
                cbs = bc.get(kind)
                if cbs is None:
                    continue
                cbs = _array(cbs, f"boundary_conditions.{kind}")
                for flag in np.unique(bf[bt == code]):
                    if flag >= len(cbs):
                        raise DeckError(f"no {kind} callback for flag {flag}")
                    sel = np.nonzero((bt == code) & (bf == flag))[0]
                    pts = xb[sel].reshape(-1, self.ndim)
                    bv[sel] = self._values(cbs[flag], pts, disc.neq, f"{kind} callback").reshape(
                        len(sel), space.nfqp, disc.neq)
            disc.set_boundary_values(torch.from_numpy(bv.ravel()).cuda())
        src = _get(self.cfg, "conservation_law", "source")
        if src is not None:
            xq = space.qp_coords().reshape(-1, self.ndim)
            sq = self._values(src, xq, disc.neq, "source")
            disc.set_source_values(torch.from_numpy(sq.ravel()).cuda())

    # ---- output
    def _writer(self, geo, u_host, neq, collection):
        w = str(_get(self.cfg, "output", "writer", default="")).lower()
        if not w:
            return None
        d = os.path.join(self.workdir, "iceicle_data")
        if w == "dat":
            if self.ndim != 1:
                raise DeckError("the dat writer is 1D")
            wr = io.DatWriter(geo, directory=d, collection_name=collection)
            wr.register_fields(u_host, neq)
            return lambda it, t: wr.write_dat(it, t)
        if w == "vtu":
            wr = io.PVDWriter(geo, directory=d, collection_name=collection)
            names = ["u"] if neq == 1 else ["rho"] + [f"rhou{d}" for d in range(self.ndim)] + ["rhoe"]
            wr.register_fields(u_host, neq, names)
            return lambda it, t: wr.write_vtu(it, t)
        raise DeckError(f"writer {w!r} is not supported")

    # ---- run
    def run(self):
        """Execute the deck; returns a dict with the final state (host array),
        the solver report and the post-processing results."""
        import torch
        self.mesh_desc = self.mesh()
        order = int(_get(self.cfg, "fespace", "order", default=1))
        if not 1 <= order <= 4:
            raise DeckError(f"fespace.order {order} is outside 1..4")
        quad = str(_get(self.cfg, "fespace", "quadrature", default="gauss")).lower()
        if quad != "gauss":
            raise DeckError(f"quadrature {quad!r} is not supported")
        basis = str(_get(self.cfg, "fespace", "basis", default="lagrange")).lower()
        if basis not in ("lagrange", "legendre"):
            raise DeckError(f"basis {basis!r} is not supported")
        params = self.params()
        space = engine.FESpace(self.mesh_desc, order)
        disc = engine.ConservationLawDDG(space, params)
        self.boundary_and_source(space, disc)
        u = torch.zeros(disc.size, dtype=torch.float64, device="cuda")
        self.initial_condition(space, disc, u)
        geo = io.HostGeometry(self.mesh_desc, order)
        u_host = u.cpu().numpy().copy()
        w0 = self._writer(geo, u_host, disc.neq, "initial_condition")
        if w0:
            w0(0, 0.0)
        result = {"space": space, "disc": disc, "u": u}
        result["solver"] = self._solve(space, disc, u, geo)
        result["post"] = self._post(space, disc, u, geo)
        result["state"] = u.cpu().numpy()
        return result

    def _solve(self, space, disc, u, geo):
        import torch
        sp = self.cfg.get("solver")
        if sp is None:
            return None
        kind = str(sp.get("type")).lower()
        if kind in ("explicit_euler", "rk3-ssp", "rk3-tvd"):
            dt, cfl = sp.get("dt"), sp.get("cfl")
            tf, nt = sp.get("tfinal"), sp.get("ntime")
            if (dt is None) == (cfl is None):
                raise DeckError("an explicit solver needs exactly one of dt and cfl")
            if (tf is None) == (nt is None):
                raise DeckError("an explicit solver needs exactly one of tfinal and ntime")
            k = {"explicit_euler": abi.STEP_EXPLICIT_EULER, "rk3-ssp": abi.STEP_RK3_SSP,
                 "rk3-tvd": abi.STEP_RK3_TVD}[kind]
            desc = abi.stepper_desc(k, reference_compat=False, dt=None if dt is None else float(dt),
                                    cfl=None if cfl is None else float(cfl),
                                    tfinal=None if tf is None else float(tf), ntime=0 if nt is None else int(nt))
            st = engine.ExplicitStepper(disc, desc)
            ivis = int(sp.get("ivis") or 1)
            host = u.cpu().numpy().copy()
            wr = self._writer(geo, host, disc.neq, "data")
            r = torch.empty_like(u)
            steps = 0
            while True:
                status = st.status()
                if status["done"]:
                    break
                st.step(u)
                steps += 1
                if steps % ivis == 0:
                    engine.form_residual(space, disc, u, r)
                    s = st.status()
                    self.out(f"itime: {s['itime']:6d} | t: {s['time']:14.8g} | residual l2: "
                             f"{float(torch.linalg.norm(r)):14.8g}")
                    if wr:
                        host[:] = u.cpu().numpy()
                        wr(s["itime"], s["time"])
            s = st.status()
            return {"steps": s["itime"], "time": s["time"]}
        if kind == "newton":
            ls = sp.get("linesearch")
            lstype = str(_get(ls, "type", default="none")).lower() if ls is not None else "none"
            lsk = {"none": abi.LS_NONE, "wolfe": abi.LS_WOLFE, "cubic": abi.LS_WOLFE,
                   "corrigan": abi.LS_CORRIGAN}.get(lstype)
            if lsk is None:
                raise DeckError(f"linesearch {lstype!r} is not supported")
            kw = dict(tau_abs=float(sp.get("tau_abs") if sp.get("tau_abs") is not None else np.finfo(float).eps),
                      tau_rel=float(sp.get("tau_rel") or 0.0), kmax=int(sp.get("kmax") or 5), linesearch=lsk)
            if ls is not None:
                for key, attr in (("kmax", "ls_kmax"), ("alpha_initial", "ls_alpha_initial"),
                                  ("alpha_max", "ls_alpha_max")):
                    v = ls.get(key)
                    if v is not None:
                        kw[attr] = float(v) if attr != "ls_kmax" else int(v)
            newton = engine.PetscNewton(disc, **kw)
            rep = newton.solve(u)
            self.out(f"Newton: {rep}")
            return rep
        raise DeckError(f"solver type {kind!r} is not supported")

    def _post(self, space, disc, u, geo):
        post = self.cfg.get("post")
        if post is None:
            return {}
        tasks = [str(t).lower() for t in _array(post.get("tasks") or lua.LuaTable(), "post.tasks")]
        exact = post.get("exact_solution")
        out = {}
        uh = u.cpu().numpy()
        for t in tasks:
            if t == "ic_residual":
                e = engine.interface_error(disc, u)
                out[t] = e
                self.out(f"IC residual norm: {e:.12g}")
                continue
            if exact is None:
                continue
            if t in ("l2_error", "l1_error"):
                e = host_error(geo, uh, disc.neq, exact, "l2" if t == "l2_error" else "l1")
                out[t] = e
                self.out(f"{'L2' if t == 'l2_error' else 'L1'} error: {e:.12g}")
            elif t == "linf_error":
                e = host_linf_error(geo, uh, disc.neq, exact)
                out[t] = e
                self.out("L_infty error: " + " ".join(f"{v:.12g}" for v in e))
            elif t == "plot_exact_projection":
                import torch
                ue = torch.zeros_like(u)
                xq = space.qp_coords().reshape(-1, self.ndim)
                vq = self._values(exact, xq, disc.neq, "exact_solution")
                engine.project_values(disc, torch.from_numpy(vq.ravel()).to(u.device), ue)
                host = ue.cpu().numpy().copy()
                wr = self._writer(geo, host, disc.neq, "exact_solution")
                if wr:
                    wr(0, 0.0)
                out[t] = host
            else:
                raise DeckError(f"post task {t!r} is not supported")
        return out


def _gauss(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def _solution_at(geo, u, neq, xi):
    """u_h and the physical points at reference points xi [npts][ndim] of
    every element: (uq [nelem][npts][neq], xq [nelem][npts][ndim], detJ
    [nelem][npts])."""
    return geo.evaluate(u, neq, xi), geo.transform(xi), geo.det_jacobian(xi)


def host_error(geo, u, neq, exact, norm):
    """l2_error (sqrt of sum_q |u_h - f|^2 w |detJ| over elements and
    equations, l2_error.hpp:37-94) or l1_error (sum_q |u_h - f| w |detJ|,
    l2_error.hpp:120-171) at the (p+1)^d Gauss points."""
    n = geo.order + 1
    x1, w1 = _gauss(n)
    grids = np.meshgrid(*([x1] * geo.ndim), indexing="ij")
    xi = np.stack([g.ravel() for g in grids], axis=1)
    wq = np.ones(xi.shape[0])
    wg = np.meshgrid(*([w1] * geo.ndim), indexing="ij")
    for g in wg:
        wq = wq * g.ravel()
    uq, xq, detJ = _solution_at(geo, u, neq, xi)
    pts = xq.reshape(-1, geo.ndim)
    fe = lua.vectorize(exact, neq, *[pts[:, d] for d in range(geo.ndim)]).reshape(uq.shape)
    diff = uq - fe
    wd = (wq[None, :] * np.abs(detJ))[..., None]
    if norm == "l2":
        return float(np.sqrt(np.sum(diff ** 2 * wd)))
    return float(np.sum(np.abs(diff) * wd))


def host_linf_error(geo, u, neq, exact):
    """linf_error (l2_error.hpp:199-250): per equation the largest |u_h - f|
    over 101 equispaced samples per element (1D), or over the Gauss points
    (2D, 3D)."""
    if geo.ndim == 1:
        xi = (-1.0 + 0.02 * np.arange(101))[:, None]
    else:
        x1, _ = _gauss(geo.order + 1)
        grids = np.meshgrid(*([x1] * geo.ndim), indexing="ij")
        xi = np.stack([g.ravel() for g in grids], axis=1)
    uq, xq, _ = _solution_at(geo, u, neq, xi)
    pts = xq.reshape(-1, geo.ndim)
    fe = lua.vectorize(exact, neq, *[pts[:, d] for d in range(geo.ndim)]).reshape(uq.shape)
    return [float(v) for v in np.max(np.abs(uq - fe), axis=(0, 1))]


def run(path, out=print):
    """Load and run a deck file."""
    return Deck(path=path, out=out).run()
