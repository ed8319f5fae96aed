"""Every Backend method on the GPU against the oracle's CPU backend, on scaled
problems and deterministic pseudo-random vectors.

Mirrors crates/saddlepoint-cuda/tests/ops.rs:29-146: the same xorshift
vectors (every fifth entry zero so that every sign branch is taken), the same
instances (afiro, israel, 25fv47, 80bau3b, perold), and the same bars: vectors
max|a-b| / (1 + max|a|) < 1e-12, scalars |a-b| / (1 + max(|a|,|b|)) < 1e-11,
copy bitwise. Added: the benchmark generators' shapes, including rows long
enough to be split across workgroups.
"""
import numpy as np
import pytest

from conftest import netlib
from oracle import oracle as O
from saddlepoint_amd import generate
from saddlepoint_amd.backend import ScaledProblem

pytestmark = pytest.mark.gpu

MASK = (1 << 64) - 1


def vector(length: int, seed: int) -> np.ndarray:
    """ops.rs:29-39: xorshift64, values in [-2, 2], every fifth one zero."""
    s = ((seed * 0x9E3779B97F4A7C15) & MASK) | 1
    out = np.zeros(length)
    for i in range(length):
        s ^= (s << 13) & MASK
        s ^= s >> 7
        s ^= (s << 17) & MASK
        out[i] = 0.0 if i % 5 == 0 else (s >> 11) / float(1 << 53) * 4.0 - 2.0
    return out


def close(what, a, b):
    err = abs(a - b) / (1.0 + max(abs(a), abs(b)))
    assert err < 1e-11, f"{what}: cpu {a} vs gpu {b} (relative {err:.2e})"


def close_vec(what, a, b):
    assert a.shape == b.shape, f"{what}: length"
    scale = 1.0 + (np.abs(a).max() if a.size else 0.0)
    err = (np.abs(a - b).max() if a.size else 0.0) / scale
    assert err < 1e-12, f"{what}: max relative difference {err:.2e}"


def scaled_pair(p):
    """The oracle's ScaledProblem, and the same arrays as the GPU's load input."""
    so = O.ScaledProblem.of(p)
    d = so.arrays()
    sg = ScaledProblem(a=d["a"], at=d["at"], c=d["c"], lc=d["lc"], uc=d["uc"], lv=d["lv"], uv=d["uv"],
                       row_scale=d["row_scale"], col_scale=d["col_scale"], bound_scale=d["bound_scale"],
                       obj_scale=d["obj_scale"])
    return so, sg


def compare(name, p, gpu, tuning=None, panels=0):
    print(f"ops parity: {name} m={p.nrows} n={p.ncols} nnz={p.a.nnz} plan={tuning} panels={panels}", flush=True)
    cpu, sp = scaled_pair(p)
    m, n = p.nrows, p.ncols
    pg = gpu.load(sp) if tuning is None else gpu.load(sp, tuning_a=tuning, tuning_at=tuning)
    if panels:
        gpu.set_panels(pg, panels)
    up, down = gpu.upload, gpu.download
    x, y, xa, ya = vector(n, 1), vector(m, 2), vector(n, 3), vector(m, 4)
    gx, gy, gxa, gya = up(x), up(y), up(xa), up(ya)

    # SpMV.
    gm, gn = up(np.zeros(m)), up(np.zeros(n))
    cm = cpu.spmv(x)
    gpu.spmv(pg, gx, gm)
    close_vec(f"{name} spmv", cm, down(gm))
    cn = cpu.spmv_t(y)
    gpu.spmv_t(pg, gy, gn)
    close_vec(f"{name} spmv_t", cn, down(gn))

    # Steps.
    tau = 0.37
    ct, cb = cpu.primal_step(x, cn, tau)
    gt, gb = up(np.zeros(n)), up(np.zeros(n))
    gpu.primal_step(pg, gx, gn, tau, gt, gb)
    close_vec(f"{name} primal_step x_t", ct, down(gt))
    close_vec(f"{name} primal_step x_bar", cb, down(gb))

    sigma = 1.9
    cyt = cpu.dual_step(y, cm, sigma)
    gyt = up(np.zeros(m))
    gpu.dual_step(pg, gy, gm, sigma, gyt)
    close_vec(f"{name} dual_step", cyt, down(gyt))

    cz = cpu.halpern_step(x, ct, xa, 0.8, 1.0)
    gz = up(x)
    gpu.halpern_step(gz, gt, gxa, 0.8, 1.0)
    close_vec(f"{name} halpern_step", cz, down(gz))

    # Vector helpers.
    close(f"{name} distance_sq", cpu.distance_sq(x, xa), gpu.distance_sq(gx, gxa))
    close(f"{name} norm_sq", cpu.norm_sq(y), gpu.norm_sq(gy))
    close(f"{name} dot", cpu.dot(y, ya), gpu.dot(gy, gya))
    cs = cpu.sub(x, xa)
    gs = up(np.zeros(n))
    gpu.sub(gx, gxa, gs)
    close_vec(f"{name} sub", cs, down(gs))
    cs = cpu.scale(cs, -1.5)
    gpu.scale(gs, -1.5)
    close_vec(f"{name} scale", cs, down(gs))

    # Stats. Use the projected iterate so that the bounds are active.
    ps = (cpu.primal_stats(ct, cm), gpu.primal_stats(pg, gt, gm))
    close(f"{name} primal residual", ps[0].residual_sq, ps[1].residual_sq)
    close(f"{name} primal objective", ps[0].objective, ps[1].objective)
    ds = (cpu.dual_stats(y, cn), gpu.dual_stats(pg, gy, gn))
    close(f"{name} dual residual", ds[0].residual_sq, ds[1].residual_sq)
    close(f"{name} dual objective", ds[0].objective, ds[1].objective)
    pr = (cpu.primal_ray_stats(x, cm), gpu.primal_ray_stats(pg, gx, gm))
    close(f"{name} primal ray objective", pr[0].objective, pr[1].objective)
    close(f"{name} primal ray violation", pr[0].violation_sq, pr[1].violation_sq)
    close(f"{name} primal ray norm", pr[0].norm_sq, pr[1].norm_sq)
    dr = (cpu.dual_ray_stats(y, cn), gpu.dual_ray_stats(pg, gy, gn))
    close(f"{name} dual ray objective", dr[0].objective, dr[1].objective)
    close(f"{name} dual ray violation", dr[0].violation_sq, dr[1].violation_sq)
    close(f"{name} dual ray norm", dr[0].norm_sq, dr[1].norm_sq)

    sd = (cpu.step_dual_residual_sq(x, ct, cn, cs, tau), gpu.step_dual_residual_sq(pg, gx, gt, gn, gs, tau))
    close(f"{name} step dual residual", sd[0], sd[1])

    # Copy and round trip.
    gc = up(np.zeros(n))
    gpu.copy(gx, gc)
    assert np.array_equal(down(gc), x), f"{name} copy"


@pytest.mark.parametrize("name", ["afiro", "israel", "25fv47", "80bau3b", "perold"])
def test_backend_ops_match_cpu_netlib(gpu, name):
    compare(name, netlib(name), gpu)


def test_backend_ops_match_cpu_config0(gpu):
    compare("config0", generate.config0(), gpu)


def test_backend_ops_match_cpu_power_law(gpu):
    # config1's generator at 1/10 scale: Zipf(1.3) columns and dense rows of
    # 2k-6k nonzeros, so both stream blocks and multi-workgroup rows run.
    compare("config1@0.1", generate.config1(scale=0.1), gpu)


def test_backend_ops_match_cpu_very_long_rows(gpu):
    # Two rows far longer than one workgroup's chunk (4,096 nonzeros), so they are
    # split across workgroups and combined by the finishing kernel.
    p = very_long_rows()
    assert np.diff(p.a.row_ptr).max() > 4096
    compare("very-long-rows", p, gpu)


def very_long_rows():
    rng = np.random.default_rng(21)
    m, n = 3000, 150_000
    cols = np.repeat(np.arange(n), 3)
    rows = rng.integers(2, m, cols.size)
    long_cols = np.concatenate([rng.choice(n, 140_000, replace=False), rng.choice(n, 70_000, replace=False)])
    long_rows = np.concatenate([np.zeros(140_000, np.int64), np.ones(70_000, np.int64)])
    r = np.concatenate([rows, long_rows])
    c = np.concatenate([cols, long_cols])
    a = generate._csr_from_rows_cols(m, n, r, c, generate._values(rng, r.size, -2.0, 2.0))
    return generate._finish(rng, a, eq_frac=0.6)


@pytest.mark.parametrize("shift", [8, 12])
def test_backend_ops_match_cpu_with_column_panels(gpu, shift):
    # Column panels: the rows with more than 64 nonzeros gathered panel by
    # panel (256 or 4,096 columns per panel) and their panel sums added in
    # panel order; every A' product and step still matches the CPU backend.
    compare("config1@0.1", generate.config1(scale=0.1), gpu, panels=shift)
    compare("config2@0.02", generate.config2(scale=0.02), gpu, panels=shift)


def test_backend_ops_match_cpu_with_wide_column_panels(gpu):
    # Panels of 65,536 columns over rows of 140k and 70k nonzeros: a row's
    # share of one panel spans several workgroups of the panel matrix, which
    # its own finishing pass adds before the panel sums are added.
    compare("very-long-rows", very_long_rows(), gpu, panels=16)


@pytest.mark.parametrize("tuning", [(256, 256, 256), (512, 64, 1024), (1024, 1024, 37), (2048, 100, 1024)])
def test_backend_ops_match_cpu_under_other_plans(gpu, tuning):
    # Every SpMV plan the autotuner may pick (smaller stream blocks, rows
    # split off as long rows, few rows per block) computes the same products.
    compare("israel", netlib("israel"), gpu, tuning)
    compare("config1@0.05", generate.config1(scale=0.05), gpu, tuning)


def test_spmv_is_deterministic(gpu):
    p = generate.config1(scale=0.05)
    _, sp = scaled_pair(p)
    pg = gpu.load(sp)
    x = gpu.upload(vector(p.ncols, 7))
    outs = []
    for _ in range(3):
        o = gpu.zeros(p.nrows)
        gpu.spmv(pg, x, o)
        outs.append(gpu.download(o))
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    v = gpu.upload(vector(p.nrows, 8))
    d = [gpu.dot(v, v) for _ in range(3)]
    assert d[0] == d[1] == d[2]


def test_rows_inside_one_piece_sum_left_to_right(gpu):
    # A row that falls inside one thread's merge-path piece is summed left to
    # right, exactly as the CPU backend does, so those rows are bitwise equal.
    p = generate.config0()
    cpu, sp = scaled_pair(p)
    pg = gpu.load(sp)
    x = vector(p.ncols, 11)
    out = gpu.zeros(p.nrows)
    gpu.spmv(pg, gpu.upload(x), out)
    g = gpu.download(out)
    c = cpu.spmv(x)
    # Rows that cross pieces are summed in another fixed order: 1e-12 on all,
    # bitwise on a good share.
    close_vec("spmv", c, g)
    assert np.mean(g == c) > 0.3


def test_shape_errors_are_reported(gpu):
    from saddlepoint_amd._lib import SP_ERR_SHAPE, BackendError

    p = netlib("afiro")
    _, sp = scaled_pair(p)
    pg = gpu.load(sp)
    with pytest.raises(BackendError) as e:
        gpu.spmv(pg, gpu.zeros(p.ncols + 1), gpu.zeros(p.nrows))
    assert e.value.code == SP_ERR_SHAPE
    with pytest.raises(BackendError) as e:
        gpu.dot(gpu.zeros(3), gpu.zeros(4))
    assert e.value.code == SP_ERR_SHAPE


def test_int32_limit_is_reported(gpu):
    """More rows or nonzeros than 32-bit device indices hold: SP_ERR_TOO_LARGE
    (the reference's CudaError::TooLarge, sparse.rs:60-62), before any array
    is touched."""
    import ctypes as C

    from saddlepoint_amd._lib import SP_ERR_TOO_LARGE, CsrView, ScaledView, lib

    big = 1 << 31
    v = ScaledView()
    v.a = CsrView(big, 1, 1, None, None, None)
    v.at = CsrView(1, big, 1, None, None, None)
    h = C.c_void_p()
    rc = lib().sp_problem_load(gpu.h, C.byref(v), C.byref(h))
    assert rc == SP_ERR_TOO_LARGE and not h.value
    assert "32-bit" in lib().sp_last_error(gpu.h).decode()


@pytest.mark.parametrize("kind,per", [("copy", 16), ("triad", 24), ("copy_f32", 8), ("triad_f32", 12)])
def test_bandwidth_probes_report_plausible_rates(gpu, kind, per):
    """The copy and triad probes in f64 and f32 (the counterpart of the
    reference's probe.rs:38-41) time a kernel that moves `per` bytes per
    element: a finite rate above zero and below the 8 TB/s HBM peak."""
    n = 1 << 26
    ms = gpu.probe(kind, n, reps=5)
    gbs = per * n / (ms * 1e-3) / 1e9
    print(f"{kind}: {ms:.4f} ms per launch, {gbs:.0f} GB/s", flush=True)
    assert np.isfinite(ms) and ms > 0.0
    assert 0.0 < gbs <= 8000.0


def test_overlap_probe_kinds(gpu):
    """The gather/stream overlap probe: every kind times a finite launch, the
    stream alone stays under the HBM peak, and doing both kinds of work takes
    longer than either alone (each does all of its own blocks); a ragged n is
    rounded up to whole blocks and a table past int32 indices is refused."""
    n, table = (1 << 21) + 5, 400_000
    t = {k: gpu.probe("overlap_" + k, n, table, reps=5) for k in ("gather", "stream", "alternate", "both")}
    print({k: round(v * 1e3, 2) for k, v in t.items()}, flush=True)
    assert all(np.isfinite(v) and v > 0.0 for v in t.values())
    blocks = -(-n // 2048)
    assert 12 * blocks * 2048 / (t["stream"] * 1e-3) / 1e9 <= 8000.0
    for k in ("alternate", "both"):
        assert t[k] > max(t["gather"], t["stream"])
    with pytest.raises(RuntimeError):
        gpu.probe("overlap_gather", n, 0, reps=1)
    with pytest.raises(RuntimeError):
        gpu.probe("overlap_gather", n, 1 << 31, reps=1)


def test_gather_replay_is_no_slower_than_the_spmv(gpu):
    """The gather replay (the matrix's own index stream, gathers only) is the
    ceiling of its SpMV's gathers: it takes no longer than the SpMV itself."""
    from saddlepoint_amd.autotune import structural

    pg = gpu.load(structural(generate.config1(scale=0.25)))
    for transpose in (False, True):
        replay = gpu.gather_replay_time(pg, transpose=transpose, reps=10)
        spmv = gpu.spmv_time(pg, transpose=transpose, reps=10)
        print(f"transpose={transpose}: replay {replay * 1e3:.1f} us, spmv {spmv * 1e3:.1f} us", flush=True)
        assert 0.0 < replay <= spmv
    pg.free()


@pytest.mark.parametrize("n", [0, 1, 1000, (1 << 20) + 7])
def test_device_unscale_is_the_host_unscale(gpu, n):
    """sp_unscale_primal / sp_unscale_dual (scaling.rs:114-121) are bitwise
    x_s · col_scale / bound_scale and y_s · row_scale / obj_scale as the host
    computes them (the product, then the division), over signs, zeros and
    factors from 2^-40 to 2^40 (subnormal and overflowing products included)."""
    rng = np.random.default_rng(n + 11)
    v = rng.standard_normal(n) * np.exp2(rng.integers(-60, 60, n))
    f = np.exp2(rng.uniform(-40.0, 40.0, n))
    if n:
        v[0] = 0.0
        v[-1] = -v[-1]
    if n > 3:
        v[1], f[1] = 1e-300, 1e-20    # a subnormal product
        v[2], f[2] = 1e300, 1e20      # an overflowing product: +inf
    for s in (1.0, 3.7, 2.0 ** -30):
        with np.errstate(over="ignore", under="ignore"):
            want = v * f / s
        dv, df, out = gpu.upload(v), gpu.upload(f), gpu.zeros(n)
        gpu.unscale_primal(dv, df, s, out)
        got_x = gpu.download(out)
        gpu.unscale_dual(dv, df, s, out)
        got_y = gpu.download(out)
        assert np.array_equal(got_x, want) and np.array_equal(got_y, want)
        assert np.array_equal(np.signbit(got_x), np.signbit(want))


def test_device_unscale_rejects_mismatched_lengths(gpu):
    from saddlepoint_amd._lib import BackendError

    with pytest.raises(BackendError, match="wrong length"):
        gpu.unscale_primal(gpu.upload(np.ones(5)), gpu.upload(np.ones(4)), 1.0, gpu.zeros(5))
    with pytest.raises(BackendError):
        gpu.unscale_dual(gpu.upload(np.ones(5)), gpu.upload(np.ones(5)), 1.0, gpu.zeros(6))
