"""Drivers for the oracle's residual-terminated Jacobi and its FP64 restatement of passive
networks (test infrastructure)."""
import ctypes as C

import numpy as np

import harness as H

P = C.c_void_p


def lib():
    L = H.oracle_lib()
    L.oracle_set_jacobi.argtypes = [P, C.c_double, C.c_int]
    L.oracle_fp64_create.restype = P
    L.oracle_fp64_create.argtypes = [P, C.c_uint32, P, C.c_uint32]
    L.oracle_fp64_set_jacobi.argtypes = [P, C.c_double, C.c_int]
    L.oracle_fp64_run_ticks.argtypes = [P, C.c_float, C.c_uint32]
    L.oracle_fp64_read.argtypes = [P] * 6
    L.oracle_fp64_destroy.argtypes = [P]
    return L


def fp64_run(net, ticks, dt, tol=0.0, max_sweeps=50):
    """FP64 state after `ticks` ticks: dict of V (n), intra/extra (n x 7), ATP, temperature."""
    L = lib()
    h = L.oracle_fp64_create(net.segments.ctypes.data, len(net.segments), net.connections.ctypes.data,
                             len(net.connections))
    L.oracle_fp64_set_jacobi(h, tol, max_sweeps)
    L.oracle_fp64_run_ticks(h, dt, ticks)
    n = len(net.segments)
    out = dict(V=np.zeros(n), intra=np.zeros((n, 7)), extra=np.zeros((n, 7)), atp=np.zeros(n), temp=np.zeros(n))
    L.oracle_fp64_read(h, out["V"].ctypes.data, out["intra"].ctypes.data, out["extra"].ctypes.data,
                       out["atp"].ctypes.data, out["temp"].ctypes.data)
    L.oracle_fp64_destroy(h)
    return out


def field_errors(segments, ref):
    """Largest relative error of each field of `segments` against the FP64 state."""
    def rel(x, y):
        return float((np.abs(x - y) / np.maximum(np.abs(y), 1e-12)).max())
    err = {"V_m": rel(segments["V_m"], ref["V"]), "ATP": rel(segments["ATP_level"], ref["atp"]),
           "temperature": rel(segments["temperature"], ref["temp"])}
    for k in range(7):
        err[f"intra[{k}]"] = rel(segments["intra_ions"][:, k], ref["intra"][:, k])
        err[f"extra[{k}]"] = rel(segments["extra_ions"][:, k], ref["extra"][:, k])
    return err
