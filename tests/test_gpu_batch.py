"""Batched path (many configurations in one set of launches) and the
force-matching adjoint.

Bars: FP64 batch forces equal the single-configuration FP64 path on each
configuration to 1e-10 relative; the adjoint gradient equals central finite
differences of the FP64 loss (h = 1e-5 relative step) to 1e-6 relative; a
replayed hipGraph reproduces the eager loss and gradient; the FP32 batch
forces are within 1e-5 relative RMS of FP64 and its gradient within the
stated 1e-5 relative."""
import copy

import numpy as np
import pytest

from mdff_amd import configs as C
from mdff_amd import params as P

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _tight(s, tol=1e-13):
    p = s.params
    p.calc_mode = P.CALC_OTHER
    p.conv_tol_ind = tol
    p.max_scf_pol_iter = 400
    return s


def _systems(nconf, nmax=240, kmax=4, seed=4, tol=1e-13):
    out = C.batch(nconf=nconf, seed=seed, nmin=96, nmax=nmax, kmax=kmax, precision=P.PREC_FP64)
    return [_tight(s, tol) for s in out]


def _positions(systems):
    import torch
    r = np.concatenate([s.r for s in systems], axis=0)
    return torch.tensor(r.T.copy(), dtype=torch.float64, device="cuda")


def _forcefield_forces(s):
    from mdff_amd.field import ForceField
    ff = ForceField(s.params, s.itype, s.A)
    out = ff.engforce(s.r, itime=1)
    ff.close()
    return out["f"]


def test_batch_matches_single_configuration_path():
    from mdff_amd.batch import Batch
    systems = _systems(3)
    b = Batch.from_systems(systems, precision=P.PREC_FP64)
    b.set_positions(_positions(systems))
    f, (ns, rmsd) = b.forces(0)
    f = f.cpu().numpy()
    assert rmsd <= 1e-13 and ns > 3
    for c, s in enumerate(systems):
        ref = _forcefield_forces(s)
        got = f[:, b.offsets[c]:b.offsets[c + 1]].T
        assert rel_rms(got, ref) <= 1e-10, (c, rel_rms(got, ref))
    b.close()


def _target(b, systems, seed=41):
    """Target forces of a nearby hidden model plus noise (the force-matching
    set-up): forces at theta * (1 + 2% noise), plus N(0, 0.05) eV/A."""
    import torch
    from mdff_amd.batch import get_theta, set_theta
    rng = np.random.default_rng(seed)
    th0 = get_theta(b.model)
    hidden = th0 * (1.0 + 0.02 * rng.standard_normal(th0.shape))
    b.set_model(set_theta(b.model, hidden))
    f, _ = b.forces(0)
    b.set_model(set_theta(b.model, th0))
    noise = torch.tensor(0.05 * rng.standard_normal(f.shape), device="cuda")
    return (f + noise).contiguous()


def test_adjoint_gradient_matches_finite_differences():
    from mdff_amd.batch import Batch, get_theta, theta_names
    systems = _systems(2, nmax=150, kmax=3)
    b = Batch.from_systems(systems, precision=P.PREC_FP64)
    b.set_positions(_positions(systems))
    target = _target(b, systems)
    th0 = get_theta(b.model)
    loss, grad = b.loss_grad(target, 0)
    assert loss > 0
    names = theta_names(b.model)
    h = 1e-5
    worst = 0.0
    for k in range(len(th0)):
        step = h * max(abs(th0[k]), 1.0)
        thp = th0.copy()
        thp[k] += step
        b.set_theta(thp)
        lp, _ = b.loss_grad(target, 0)
        thm = th0.copy()
        thm[k] -= step
        b.set_theta(thm)
        lm, _ = b.loss_grad(target, 0)
        b.set_theta(th0)
        fd = (lp - lm) / (2 * step)
        err = abs(fd - grad[k]) / max(abs(fd), 1e-12)
        print(f"{names[k]:10s} adjoint {grad[k]: .10e} fd {fd: .10e} rel {err:.2e}")
        if abs(fd) > 1e-9 * np.max(np.abs(grad)):
            worst = max(worst, err)
    assert worst <= 1e-6, worst
    b.close()


def test_adjoint_morse_and_damping_b_match_finite_differences():
    """Morse pair parameters (recomputed pair terms) and the Tang-Toennies
    damping b (explicit damped static field and force, plus the dipole
    tangent) against central differences of the FP64 loss."""
    from mdff_amd.batch import Batch, get_theta, theta_names
    systems = _systems(2, nmax=150, kmax=3)
    p = systems[0].params
    p.lmorse = 1
    for ti in range(2):
        for tj in range(2):
            k = ti + 16 * tj
            p.epsmor[k] = 0.04 + 0.01 * (ti + tj)
            p.sigmamor[k] = 1.9 + 0.2 * (ti + tj)
            p.rhomor[k] = 1.8 + 0.1 * ti * tj
    b = Batch.from_systems(systems, precision=P.PREC_FP64)
    b.set_positions(_positions(systems))
    names = theta_names(b.model)
    assert "eps_mo[12]" in names and "b_TT[2]" in names, names
    target = _target(b, systems)
    th0 = get_theta(b.model)
    loss, grad = b.loss_grad(target, 0)
    h = 1e-5
    worst = 0.0
    for k, name in enumerate(names):
        if not (name.endswith("_mo[11]") or "_mo[" in name or name.startswith("b_TT")):
            continue
        step = h * max(abs(th0[k]), 1.0)
        lv = []
        for sgn in (1.0, -1.0):
            th = th0.copy()
            th[k] += sgn * step
            b.set_theta(th)
            lv.append(b.loss_grad(target, 0)[0])
        b.set_theta(th0)
        fd = (lv[0] - lv[1]) / (2 * step)
        err = abs(fd - grad[k]) / max(abs(fd), 1e-12)
        print(f"{name:12s} adjoint {grad[k]: .10e} fd {fd: .10e} rel {err:.2e}")
        worst = max(worst, err)
    b.close()
    assert worst <= 1e-6, worst


def test_autograd_function_and_graph_replay():
    import torch
    from mdff_amd.batch import Batch, force_match_loss, get_theta
    systems = _systems(3, nmax=200, kmax=3, tol=1e-10)
    b = Batch.from_systems(systems, precision=P.PREC_FP64)
    b.set_positions(_positions(systems))
    target = _target(b, systems)
    loss, grad = b.loss_grad(target, 20)
    theta = torch.tensor(get_theta(b.model), dtype=torch.float64, requires_grad=True)
    L = force_match_loss(theta, b, target, 20)
    L.backward()
    assert abs(L.item() - loss) <= 1e-14 * abs(loss)
    assert np.allclose(theta.grad.numpy(), grad, rtol=1e-14, atol=0)
    # graph: capture on the first call, replay with unrelated eager work
    # (and synchronizations) between the replays; every replay equals the
    # eager result
    b.graph(True)
    for rep in range(3):
        lg, gg = b.loss_grad(target, 20)
        x = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        assert abs(lg - loss) <= 1e-13 * abs(loss), rep
        assert np.allclose(gg, grad, rtol=1e-12, atol=1e-14 * np.max(np.abs(grad))), rep
        del x
    b.graph(False)
    b.close()


def test_mixed_batch_against_fp64():
    from mdff_amd.batch import Batch
    systems = _systems(4, nmax=400, kmax=5, tol=1e-10)
    b64 = Batch.from_systems(systems, precision=P.PREC_FP64)
    b32 = Batch.from_systems(systems, precision=P.PREC_MIXED)
    pos = _positions(systems)
    b64.set_positions(pos)
    b32.set_positions(pos)
    target = _target(b64, systems)
    l64, g64, f64 = b64.loss_grad(target, 25, want_forces=True)
    l32, g32, f32 = b32.loss_grad(target, 25, want_forces=True)
    ferr = rel_rms(f32.cpu().numpy(), f64.cpu().numpy())
    gerr = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    print(f"mixed batch: force rel RMS {ferr:.2e}, loss rel {abs(l32 - l64) / l64:.2e}, "
          f"gradient max rel {gerr:.2e}")
    assert ferr <= 1e-5
    assert abs(l32 - l64) <= 1e-4 * l64
    assert gerr <= 1e-5
    b64.close()
    b32.close()


def test_mixed_launch_shapes_agree_and_autotune():
    """Every launch shape of the FP32 matrix-core reciprocal sums (K slices
    of the apply, atoms per structure-factor split) gives the same loss and
    gradient to FP32 summation-order noise; the autotuner keeps one of them;
    an FP64 batch has no such shape."""
    from mdff_amd.batch import MIXED_CANDIDATES, Batch, autotune
    systems = _systems(3, nmax=300, kmax=4, tol=1e-10)
    b = Batch.from_systems(systems, precision=P.PREC_MIXED)
    b.set_positions(_positions(systems))
    target = _target(b, systems)
    ref = None
    for slices, split in ((2, 2048), (1, 512), (4, 128)):
        assert b.set_mixed_tuning(slices, split)
        loss, grad = b.loss_grad(target, 20)
        if ref is None:
            ref = (loss, grad)
            continue
        assert abs(loss - ref[0]) <= 1e-6 * abs(ref[0]), (slices, split)
        assert np.allclose(grad, ref[1], rtol=1e-4, atol=1e-6 * np.max(np.abs(ref[1])))
    rec = autotune(b, target, 20, reps=1, use_cache=False)
    assert rec["best"] in [dict(c) for c in MIXED_CANDIDATES]
    b.close()
    b64 = Batch.from_systems(systems, precision=P.PREC_FP64)
    assert not b64.set_mixed_tuning(2, 2048)
    b64.close()
