"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs, FP64 bar 1e-10 per cell (metric in tests/parity.py)."""
import numpy as np
import pytest

from tests.parity import TOL_FP64, assert_parity, gaussian_ic, rest_frame, vah

pytestmark = pytest.mark.gpu


def run_pair(oracle_mod, nx, ny, nz, steps, e, u=None, pl=None, pi=None, dx=0.1, dz=0.1, dt=0.01, etabar=0.2,
             t0=0.5, flux_mode=0, eos_table=None):
    V = vah()
    n = nx * ny * nz
    ut, ux, uy, un = rest_frame(n) if u is None else u
    g = V.Vah(nx, ny, nz, dx, dx, dz, dt, etabar, t0, flux_mode=flux_mode, eos_table=eos_table)
    o = oracle_mod.Oracle(nx, ny, nz, dx, dx, dz, dt, etabar, t0, eos_table=eos_table)
    g.set_primitives(e, ut, ux, uy, un, pl=pl, pi=pi)
    o.set_primitives(e, ut, ux, uy, un, pl=pl, pi=pi)
    g.step(steps)
    g.synchronize()
    o.step(steps)
    assert abs(g.time - o.time) == 0.0
    a, b, tau = g.state(), o.state(), o.time
    g.close()
    o.close()
    return a, b, tau


@pytest.mark.parametrize("flux_mode", [0, 1], ids=["shared", "per_cell"])
def test_gaussian_81x81_200_steps(oracle_mod, flux_mode):
    """Elliptic Gaussian profile on the 81x81x1 lattice of the optical-Glauber
    event (dx 0.1, tau0 0.5, dtau 0.01, 200 steps, conformal, eta/s 0.2,
    P_L = e/3); the optical-Glauber event itself is in test_gpu_workloads."""
    e = gaussian_ic(81, 81)
    a, b, tau = run_pair(oracle_mod, 81, 81, 1, 200, e, pl=e / 3, flux_mode=flux_mode)
    worst = assert_parity(a, b, tau, label="81x81x1 Gaussian")
    print("worst field", worst)


def test_reference_pl_init_and_lumps(oracle_mod):
    """P_L = 0.0154483 e (the reference's hard-coded init), lumpy seeded IC, 60 steps."""
    e = gaussian_ic(48, 40, seed=5, tilt=0.3)
    a, b, tau = run_pair(oracle_mod, 48, 40, 1, 60, e)
    assert_parity(a, b, tau, label="lumpy")


@pytest.mark.parametrize("flux_mode", [0, 1], ids=["shared", "per_cell"])
def test_3d_small(oracle_mod, flux_mode):
    """3+1D 24x20x14 with an eta plateau, 25 steps, in both flux modes (faces
    shared between cells, and the reference's per-cell work split)."""
    e = gaussian_ic(24, 20, 14, dz=0.3, seed=9)
    a, b, tau = run_pair(oracle_mod, 24, 20, 14, 25, e, pl=e / 3, dz=0.3, dt=0.01, flux_mode=flux_mode)
    assert_parity(a, b, tau, label="3D")


@pytest.mark.parametrize("dims", [(1, 1, 1), (1, 7, 1), (5, 1, 3), (2, 2, 2), (3, 4, 1)],
                         ids=["1x1x1", "1x7x1", "5x1x3", "2x2x2", "3x4x1"])
def test_degenerate_lattices(oracle_mod, dims):
    """Lattices one or two cells wide in some direction, down to a single
    cell: a tile with a handful of live lanes, ghost copies that are the cell
    itself, 20 steps against the oracle in both flux modes."""
    nx, ny, nz = dims
    e = gaussian_ic(nx, ny, nz, sx=0.3, sy=0.4, seed=9)
    n = e.size
    o = oracle_mod.Oracle(nx, ny, nz, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.step(20)
    b, tau = o.state(), o.time
    o.close()
    V = vah()
    for mode in (V.FLUX_SHARED, V.FLUX_PER_CELL):
        g = V.Vah(nx, ny, nz, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5, flux_mode=mode)
        g.set_primitives(e, *rest_frame(n), pl=e / 3)
        g.step(20)
        g.synchronize()
        assert g.time == tau
        assert_parity(g.state(), b, tau, label=f"{nx}x{ny}x{nz} mode {mode}")
        g.close()


@pytest.mark.parametrize("dims", [(17, 31, 2), (33, 16, 4), (16, 15, 7)], ids=["17x31x2", "33x16x4", "16x15x7"])
def test_3d_ragged_tiles_and_eta_chunks(oracle_mod, dims):
    """3+1D lattices whose sizes are not multiples of the 16 x 15 tile or of the
    3-plane eta chunk (one chunk shorter than 3, a last chunk of 1 plane, a
    single tile), tilted lumpy IC, 12 steps."""
    nx, ny, nz = dims
    e = gaussian_ic(nx, ny, nz, dz=0.4, sx=0.6, sy=0.9, tilt=0.4, seed=13)
    a, b, tau = run_pair(oracle_mod, nx, ny, nz, 12, e, pl=e / 3, dz=0.4, dt=0.01)
    assert_parity(a, b, tau, label="3D ragged")


def test_flow_and_shear_initial_state(oracle_mod):
    """Initial flow and pi from an evolved oracle state, then both continue."""
    nx, ny, nz = 40, 36, 1
    e = gaussian_ic(nx, ny, seed=3)
    n = nx * ny
    o = oracle_mod.Oracle(nx, ny, nz, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.step(80)
    st = o.state()
    tau0 = o.time
    V = vah()
    g = V.Vah(nx, ny, nz, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    g.set_state(80, st["q"], st["e"], st["p"], st["u"], st["up"])
    assert g.time == tau0
    g.step(40)
    g.synchronize()
    o.step(40)
    assert_parity(g.state(), o.state(), o.time, label="flow")
    assert np.max(np.abs(st["u"][1])) > 1e-3  # the continued state really flows


def test_ghost_cells_match_neighbours(oracle_mod):
    """Zero-gradient ghosts: a state uniform in y evolves identically in every row."""
    nx, ny = 30, 9
    x = (np.arange(nx) - (nx - 1) / 2) * 0.1
    e = np.tile(50.0 * np.exp(-x ** 2 / 2.0) + 1e-3, ny)
    a, b, tau = run_pair(oracle_mod, nx, ny, 1, 30, e, pl=e / 3)
    E = a["e"].reshape(ny, nx)
    assert np.all(E == E[0])  # no y structure may appear
    assert_parity(a, b, tau, label="ghosts")


def test_lqcd_table_eos_3d(oracle_mod):
    """The tabulated crossover EOS (10^4 nodes, linear in ln e) on a 3+1D lattice."""
    V = vah()
    table = V.eos_lqcd_table(10000, 1e-8, 1e4)
    e = gaussian_ic(20, 18, 10, dz=0.3, seed=4, peak=60.0)
    a, b, tau = run_pair(oracle_mod, 20, 18, 10, 20, e, pl=e / 3, dz=0.3, dt=0.02, etabar=0.16, eos_table=table)
    assert_parity(a, b, tau, label="table EOS")


def test_validity_diagnostics_3d(oracle_mod):
    """Knudsen and inverse Reynolds maps (first and second order) after 15
    steps of a flowing 3+1D state: the GPU's validity diagnostics against the
    oracle's, per cell within 1e-10 of each map's largest value, and R_pi2
    per cell within 1e-9 of its own value."""
    V = vah()
    nx, ny, nz = 22, 20, 9
    e = gaussian_ic(nx, ny, nz, dz=0.3, seed=12, peak=80.0)
    n = nx * ny * nz
    g = V.Vah(nx, ny, nz, 0.1, 0.1, 0.3, 0.01, 0.2, 0.5)
    o = oracle_mod.Oracle(nx, ny, nz, 0.1, 0.1, 0.3, 0.01, 0.2, 0.5)
    g.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    g.step(15)
    g.synchronize()
    o.step(15)
    a, b = g.validity(), o.validity()
    for k in ("kn_T", "kn_L", "kn", "r_pi", "r_pi2"):
        scale = np.max(np.abs(b[k]))
        err = np.max(np.abs(a[k] - b[k])) / scale
        print(k, "max", scale, "worst", err)
        assert scale > 0 and err < 1e-10, k
    rel = np.abs(a["r_pi2"] / b["r_pi2"] - 1)
    print("r_pi2 per-cell relative: max", rel.max(), "median", np.median(rel))
    assert rel.max() < 1e-9
    g.close()
    o.close()


def test_wuppertal_budapest_eos_3d(oracle_mod):
    """The reference's non-conformal EOS (Wuppertal-Budapest fits) on a 3+1D lattice."""
    V = vah()
    nx, ny, nz, steps = 20, 18, 10, 20
    e = gaussian_ic(nx, ny, nz, dz=0.3, seed=6, peak=60.0)
    n = nx * ny * nz
    g = V.Vah(nx, ny, nz, 0.1, 0.1, 0.3, 0.02, 0.16, 0.5, eos_kind=V.EOS_WB)
    o = oracle_mod.Oracle(nx, ny, nz, 0.1, 0.1, 0.3, 0.02, 0.16, 0.5, eos_wb=True)
    g.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    g.step(steps)
    g.synchronize()
    o.step(steps)
    assert_parity(g.state(), o.state(), o.time, label="WB EOS")
    g.close()
    o.close()


def precision_errors(a, b):
    """Per-field max and RMS relative errors of the split-precision run a
    against the FP64 reference b (e and u^mu relative to u^tau, pi to e)."""
    out = {}
    e = b["e"]
    out["e"] = np.abs(a["e"] - e) / e
    for d, name in enumerate(("ut", "ux", "uy", "un")):
        out[name] = np.abs(a["u"][d] - b["u"][d]) / b["u"][0]
    out["pl"] = np.abs(a["q"][4] - b["q"][4]) / b["q"][4]
    out["pi"] = np.max(np.abs(a["q"][5:] - b["q"][5:]), axis=0) / e
    return {k: (float(v.max()), float(np.sqrt(np.mean(v * v)))) for k, v in out.items()}


def test_fp32_split_precision_form(oracle_mod):
    """The 81x81x1 optical Glauber event with seeded +-5% lognormal noise on e, run to tau = 5 fm/c
    (450 steps) with stencil, fluxes and sources in float: e and u^mu within
    1e-4 of the FP64 oracle, total energy within 1e-6."""
    V = vah()
    nx = ny = 81
    n = nx * ny
    g0 = V.glauber()
    e = V.ic_optical_glauber(nx, ny, 1, 0.1, 0.1, g0)
    rng = np.random.default_rng(2024)
    e = e * np.exp(0.05 * rng.standard_normal(n))
    ut, ux, uy, un = rest_frame(n)
    g = V.Vah(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5, precision=V.FP32_SPLIT)
    o = oracle_mod.Oracle(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    g.set_primitives(e, ut, ux, uy, un, pl=e / 3)
    o.set_primitives(e, ut, ux, uy, un, pl=e / 3)
    g.step(450)
    g.synchronize()
    o.step(450)
    assert o.time == pytest.approx(5.0)
    a, b = g.state(), o.state()
    errs = precision_errors(a, b)
    print("split precision vs FP64 oracle (max, rms):", errs)
    E_a = np.sum(a["q"][0]) * o.time
    E_b = np.sum(b["q"][0]) * o.time
    print("total energy drift", abs(E_a / E_b - 1))
    assert errs["e"][0] < 1e-4
    for k in ("ut", "ux", "uy", "un"):
        assert errs[k][0] < 1e-4, k
    assert abs(E_a / E_b - 1) < 1e-6
    g.close()
    o.close()


def test_fp32_split_3d_against_fp64_oracle(oracle_mod):
    """The split-precision form on a 3+1D lattice (eta faces from the
    lane-private plane ring): e and u^mu within 1e-4 of the FP64 oracle after
    40 steps, in the single-lattice path and in the batch."""
    V = vah()
    nx, ny, nz, steps = 26, 22, 13, 40
    e = gaussian_ic(nx, ny, nz, dz=0.3, seed=21, peak=70.0)
    n = nx * ny * nz
    o = oracle_mod.Oracle(nx, ny, nz, 0.1, 0.1, 0.3, 0.01, 0.2, 0.5)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.step(steps)
    b = o.state()
    o.close()
    g = V.Vah(nx, ny, nz, 0.1, 0.1, 0.3, 0.01, 0.2, 0.5, precision=V.FP32_SPLIT, eta_chunk=3)
    g.set_primitives(e, *rest_frame(n), pl=e / 3)
    g.step(steps)
    g.synchronize()
    a = g.state()
    g.close()
    B = V.Batch(budget_bytes=1 << 28, max_events=2, precision=V.FP32_SPLIT, freezeout_T_GeV=0.0, steps_per_graph=0)
    s = B.admit(nx, ny, nz, 0.1, 0.1, 0.3, 0.01, 0.5, 0.2, e, pl=e / 3, max_steps=steps)
    B.step(steps)
    c = B.state(s)
    B.close()
    for label, x in (("single", a), ("batch", c)):
        errs = precision_errors(x, b)
        print(label, errs)
        assert errs["e"][0] < 1e-4, label
        for k in ("ut", "ux", "uy", "un"):
            assert errs[k][0] < 1e-4, (label, k)


@pytest.mark.parametrize("dims", [(17, 31, 2), (33, 16, 4)], ids=["17x31x2", "33x16x4"])
def test_fp32_split_3d_ragged(oracle_mod, dims):
    """The split-precision eta ring on lattices with a chunk shorter than 3
    planes and a last chunk of one plane: e and u^mu within 1e-4 of the FP64
    oracle after 20 steps."""
    V = vah()
    nx, ny, nz = dims
    steps = 20
    e = gaussian_ic(nx, ny, nz, dz=0.4, sx=0.6, sy=0.9, tilt=0.4, seed=17)
    n = nx * ny * nz
    o = oracle_mod.Oracle(nx, ny, nz, 0.1, 0.1, 0.4, 0.01, 0.2, 0.5)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.step(steps)
    b = o.state()
    o.close()
    g = V.Vah(nx, ny, nz, 0.1, 0.1, 0.4, 0.01, 0.2, 0.5, precision=V.FP32_SPLIT)
    g.set_primitives(e, *rest_frame(n), pl=e / 3)
    g.step(steps)
    g.synchronize()
    a = g.state()
    g.close()
    errs = precision_errors(a, b)
    print(dims, errs)
    assert errs["e"][0] < 1e-4
    for k in ("ut", "ux", "uy", "un"):
        assert errs[k][0] < 1e-4, k
