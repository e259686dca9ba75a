"""Adjoint gradient of the self-consistent 2D run (mean kinetic energy after
N steps of Pic::advance with a solve every step) with respect to electrode
voltages, against FP64 central finite differences of the oracle's run;
without collisions, and with MCC collisions under common random numbers
through the torch.autograd wrapper (mag2d_amd.autograd)."""
import ctypes as C
import math

import numpy as np
import pytest

from mag2d_amd import abi
from oracle import oracle as O

QE, KB = abi.QE, abi.KB
M_E = 9.11e-31


def trap(n=48, L=2e-2):
    """Grounded frame cut into 8 electrode groups (each wall in two halves)."""
    dx = L / (n - 1)
    g = abi.grid2d(n, n, dx, dx)
    mask, volt = abi.grid_build(g, abi.GEO_EMPTY, L, L, extern_field=0.0)
    group = np.full(n * n, -1, np.int32)
    G = group.reshape(n, n)
    h = n // 2
    G[0, :h], G[0, h:] = 0, 1           # x = 0 wall
    G[-1, :h], G[-1, h:] = 2, 3         # x = L wall
    G[1:-1, 0][: h - 1], G[1:-1, 0][h - 1:] = 4, 5   # z = 0 wall (corners stay with the x walls)
    G[1:-1, -1][: h - 1], G[1:-1, -1][h - 1:] = 6, 7
    return g, mask, volt, group, dx


def electrons(N, L, rng, T=11604.5, r=5e-3):
    ang, rad = rng.uniform(0, 2 * np.pi, N), r * np.sqrt(rng.uniform(0, 1, N))
    x = (L / 2 + rad * np.cos(ang)).astype(np.float32)
    z = (L / 2 + rad * np.sin(ang)).astype(np.float32)
    v = (rng.normal(size=(3, N)) * math.sqrt(KB * T / M_E)).astype(np.float32)
    return x, z, v


def oracle_J(g_o, mask, volt, parts, dV, steps, Bz, dt, mcc=None):
    """The run in FP64; with mcc, the collisions draw the device's Philox
    numbers (seed 1234, stream 0), so runs at nearby voltages share them."""
    n = g_o.nx
    u = np.zeros(n * n)
    of = O.make_fields(g_o, mask, u, x_max=(n - 1) * g_o.dx, z_max=(n - 1) * g_o.dz, geometry_empty=1,
                       selfconsistent=1, Bz=Bz)
    p = parts.copy()
    s = (O.make_species(-QE, M_E, dt, lifetime=mcc.lifetime, mcc=mcc, seed=1234, stream=0) if mcc is not None
         else O.make_species(-QE, M_E, dt))
    v = np.ascontiguousarray(volt, np.float64)
    return O.lib().o_selfconsistent_run(C.byref(of), C.byref(s), O.dptr(p), len(p), O.dptr(v), dV, steps)


def test_oracle_selfconsistent_run_equals_step_by_step_composition():
    n, L, N, steps, dt = 24, 1e-2, 300, 6, 1e-11
    g, mask, volt, group, dx = trap(n, L)
    volt = volt + np.where(group >= 0, 0.7 * (group % 3) - 0.5, 0.0)
    rng = np.random.default_rng(0)
    x, z, v = electrons(N, L, rng, r=2e-3)
    p = O.particles(N)
    p["x"], p["z"], p["vx"], p["vy"], p["vz"] = x, z, v[0], v[1], v[2]
    og = O.grid2(n, n, dx, dx)
    dV = 1e-12
    J = oracle_J(og, mask, volt, p, dV, steps, 0.1, dt)
    # the same sequence from the single-solve oracle functions
    u = np.zeros(n * n)
    of = O.make_fields(og, mask, u, x_max=L, z_max=L, geometry_empty=1, selfconsistent=1, Bz=0.1)
    u = of._keep[1]
    rho = np.zeros(n * n)
    s = O.make_species(-QE, M_E, dt, rho=rho)
    q = p.copy()
    O.lib().o_accumulate(C.byref(of), C.byref(s), O.dptr(q), N)
    for k in range(steps):
        b = rho.copy()
        O.lib().o_poisson_rhs(C.byref(og), O.dptr(mask), O.dptr(volt), O.dptr(b), O.SOLVE_STATIC, dV)
        out = np.zeros_like(b)
        assert O.lib().o_poisson_solve(C.byref(og), O.dptr(mask), O.dptr(b), O.dptr(out)) == 0
        u[:] = out
        rho[:] = 0
        s.niter = k
        O.lib().o_advance_boris(C.byref(of), C.byref(s), O.dptr(q), N, None)
        O.lib().o_advance_boundary(C.byref(of), C.byref(s), O.dptr(q), N, None)
    assert J == pytest.approx(O.lib().o_mean_energy_eV(O.dptr(q), N, M_E), rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("K", [100, 7])
def test_adjoint_gradient_matches_fp64_finite_differences(ctx, K):
    n, L, N, steps, dt, Bz = 48, 2e-2, 2000, 60, 1e-11, 0.1
    g, mask, volt, group, dx = trap(n, L)
    rng = np.random.default_rng(1)
    V = rng.uniform(-5, 5, 8)
    volt = volt + np.where(group >= 0, V[np.maximum(group, 0)], 0.0)
    x, z, v = electrons(N, L, rng)
    dV = (N / 1e12) / ((n - 1) ** 2)
    params = abi.FieldsParams(g, L, L, 1, 1, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0, Bz, 0.0, dV)
    fields = abi.Fields2D(ctx, params, mask, volt)
    P = abi.ParticlesSoA.from_host(ctx, x, z, v[0], v[1], v[2])
    sp = abi.Species2D(-QE, M_E, dt, float("inf"), 0, 0, 0)
    zero = np.zeros(n * n)
    fields.upload(abi.FIELD_U, zero)  # the solves warm-start from u: start both runs alike
    # the adjoint solves at 1e-8, as the benchmark runs them (the forward run solves at 1e-10)
    J, grad, info = abi.adjoint_energy(ctx, fields, sp, P, steps, group=group, n_groups=8, checkpoint_every=K,
                                       adjoint_tol=1e-8)
    assert ctx.status(clear=True) == 0
    # the store ends as a plain forward run leaves it
    P2 = abi.ParticlesSoA.from_host(ctx, x, z, v[0], v[1], v[2])
    fields.upload(abi.FIELD_U, zero)
    J2, _, _ = abi.adjoint_energy(ctx, fields, sp, P2, steps)
    assert J2 == J
    assert np.array_equal(P.download()["x"].view(np.uint32), P2.download()["x"].view(np.uint32))
    # FP64 oracle: J and central differences, h = 1e-4 V
    og = O.grid2(n, n, dx, dx)
    p = O.particles(N)
    p["x"], p["z"], p["vx"], p["vy"], p["vz"] = x, z, v[0], v[1], v[2]
    J64 = oracle_J(og, mask, volt, p, dV, steps, Bz, dt)
    assert J == pytest.approx(J64, rel=1e-5)
    h = 1e-4
    fd = np.zeros(8)
    for k in range(8):
        e = np.where(group == k, h, 0.0)
        fd[k] = (oracle_J(og, mask, volt + e, p, dV, steps, Bz, dt) -
                 oracle_J(og, mask, volt - e, p, dV, steps, Bz, dt)) / (2 * h)
    scale = np.abs(fd).max()
    assert scale > 0
    assert np.abs(grad - fd).max() <= 1e-4 * scale, (grad, fd)


M_HE = 4.002602 * 1.66053906660e-27


@pytest.mark.gpu
def test_adjoint_with_collisions_through_autograd_matches_fp64_finite_differences(ctx):
    """Electrons in a He background: ELASTIC and SUPERELASTIC (0.1 eV) at
    constant rates, about 1% of the electrons colliding each step (45% at
    least once over the run), so the interaction each trial picks does not
    depend on the speed.  J and its gradient come from the autograd wrapper;
    the oracle's runs at V +- h draw the same numbers, which makes its central
    differences the derivative the adjoint computes."""
    import torch
    from mag2d_amd.autograd import ElectrodeRun, electrode_energy

    n, L, N, steps, dt, Bz = 48, 2e-2, 2000, 60, 1e-11, 0.1
    g, mask, volt, group, dx = trap(n, L)
    rng = np.random.default_rng(2)
    V0 = rng.uniform(-5, 5, 8)
    x, z, v = electrons(N, L, rng)
    dV = (N / 1e12) / ((n - 1) ** 2)
    params = abi.FieldsParams(g, L, L, 1, 1, 0, abi.BC_FREE, 0.0, 0, 0.0, 0.0, 0.0, 0.0, Bz, 0.0, dV)
    fields = abi.Fields2D(ctx, params, mask, volt)
    dens, k_el, k_se, DE = 1e21, 7e-13, 3e-13, 0.1
    mcc = abi.Mcc(ctx, M_E, [abi.Target(M_HE, dens, 300.0, 0.0, abi.Particles2D())],
                  [abi.Interaction(abi.ELASTIC, 0, 0.0, k_el, 0.0, 0, 0),
                   abi.Interaction(abi.SUPERELASTIC, 0, DE, k_se, 0.0, 0, 0)], 1.0)
    omcc = O.make_mcc(M_E, 1.0, [dict(mass=M_HE, density=dens, temperature=300.0,
                                      inters=[O.inter(O.ELASTIC, M_E, M_HE, rate=k_el),
                                              O.inter(O.SUPERELASTIC, M_E, M_HE, rate=k_se, DE=DE)])])
    assert mcc.lifetime == pytest.approx(omcc.lifetime, rel=1e-14)
    assert 1 - math.exp(-dt / mcc.lifetime) == pytest.approx(0.00995, rel=1e-3)
    sp = abi.Species2D(-QE, M_E, dt, mcc.lifetime, 0, 0, 0)
    run = ElectrodeRun(ctx, fields, volt, group, 8, sp, (x, z, v[0], v[1], v[2]), steps, mcc=mcc,
                       checkpoint_every=7, adjoint_tol=1e-8)
    V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
    J = electrode_energy(V, run)
    J.backward()
    assert ctx.status(clear=True) == 0
    # the wrapper's gradient is the library's (the reverse sweep sums the
    # gather transpose with FP64 atomics: equal to rounding, not bit for bit)
    J2, g2 = run.evaluate(V0)
    assert J2 == J.item()
    np.testing.assert_allclose(g2, V.grad.numpy(), rtol=1e-7, atol=1e-9 * np.abs(g2).max())
    # FP64 oracle with the same draws
    og = O.grid2(n, n, dx, dx)
    p = O.particles(N)
    p["x"], p["z"], p["vx"], p["vy"], p["vz"] = x, z, v[0], v[1], v[2]
    volt0 = volt + np.where(group >= 0, V0[np.maximum(group, 0)], 0.0)
    J64 = oracle_J(og, mask, volt0, p, dV, steps, Bz, dt, omcc)
    assert J.item() == pytest.approx(J64, rel=1e-5)
    h = 1e-4
    fd, fd0 = np.zeros(8), np.zeros(8)
    for k in range(8):
        e = np.where(group == k, h, 0.0)
        fd[k] = (oracle_J(og, mask, volt0 + e, p, dV, steps, Bz, dt, omcc) -
                 oracle_J(og, mask, volt0 - e, p, dV, steps, Bz, dt, omcc)) / (2 * h)
        fd0[k] = (oracle_J(og, mask, volt0 + e, p, dV, steps, Bz, dt) -
                  oracle_J(og, mask, volt0 - e, p, dV, steps, Bz, dt)) / (2 * h)
    scale = np.abs(fd).max()
    assert scale > 0
    assert np.abs(V.grad.numpy() - fd).max() <= 1e-4 * scale, (V.grad.numpy(), fd)
    # the collisions matter: without them the gradient is another one
    assert np.abs(fd - fd0).max() > 0.05 * scale, (fd, fd0)
