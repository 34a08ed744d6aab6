"""Shared helpers of the test suite: tolerances and comparisons."""
import numpy as np

# Parity bar of the plain FP64 path against the oracle (relative to the
# largest magnitude of the compared quantity).
FP64_RTOL = 1e-10


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(initial=0.0), 1e-300)
    return float(np.abs(a - b).max(initial=0.0) / scale)


def assert_close(a, b, rtol=FP64_RTOL, what=""):
    e = rel_err(a, b)
    assert e <= rtol, f"{what}: relative error {e:.3e} > {rtol:.1e}"


ENERGY_KEYS = ("u_lj", "vir_lj", "u_morse", "vir_morse", "u_bmhft", "vir_bmhft", "u_coul",
               "u_pol", "pvirial_coul")


def compare_engforce(g, o, rtol=FP64_RTOL, efg=True):
    for k in ENERGY_KEYS:
        assert_close(g[k], o[k], rtol, k)
    assert_close(g["tau_nonb"], o["tau_nonb"], rtol, "tau_nonb")
    assert_close(g["tau_coul"], o["tau_coul"], rtol, "tau_coul")
    assert g["n_scf"] == o["n_scf"], (g["n_scf"], o["n_scf"])
    assert_close(g["f"], o["f"], rtol, "forces")
    assert_close(g["mu"], o["mu"], rtol, "mu")
    assert_close(g["ef"], o["ef"], rtol, "ef")
    if efg:
        assert_close(g["efg"], o["efg"], rtol, "efg")
