"""MI355X-native DG residual/Jacobian engine for ICEicle conservation laws."""
