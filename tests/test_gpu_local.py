"""GPU parity of the narrow-band entry points (localization and *LOCAL).

Every routine is called through the C ABI of the HIP library, with device
arrays (index lists and narrow_band included; n_lo / n_hi on the host), and
through the oracle with the same seeded arguments (tests/lsm_local_cases.py).

Bars:
  index lists, narrow_band, n_lo / n_hi, index_outer, integer outputs:
          bit for bit (the band builders reproduce the reference's scan and
          first-discovery order exactly);
  real outputs: double 1e-12, float 2e-5 relative to the largest magnitude
          (the float kernels compute in FP32 throughout, the reference's float
          build promotes its double literals); the average gradient (a sum
          accumulated in FP64 here, in list order in the oracle) 1e-12 / 1e-5.
"""
import numpy as np
import pytest

from lsm_cases import grow
from lsm_local_cases import BETA, GAMMA, GB, LEVEL, MARK_D1, MARK_D2, MARK_FB, MARK_GB, local_cases_for, sphere
from lsmlib_amd import toolbox
from lsmlib_amd._native import LsmError
from lsmlib_amd.f77 import box_shape
from lsmlib_amd.signatures import LOCAL, SIGNATURES
from oracle.oracle import oracle_lib

TOL = {"double": 1e-12, "float": 2e-5}
SUM_TOL = {"double": 1e-12, "float": 1e-5}


def to_device(name, args, device):
    """The case's arguments with every array but n_lo / n_hi on the device."""
    import torch
    out, it = [], iter(args)
    for k in SIGNATURES[name]:
        if k in "OK":
            continue
        a = next(it)
        if k in "AXU" and isinstance(a, np.ndarray):
            out.append(torch.from_numpy(a.copy()).to(device))
        elif k == "N":
            out.append(a.copy())
        else:
            out.append(a)
    rest = list(it)
    assert not rest
    return out


def host_of(a):
    return a.cpu().numpy() if hasattr(a, "cpu") else a


def compare(name, real, want_args, want_ret, got_args, got_ret):
    for i, (w, g) in enumerate(zip(want_args, got_args)):
        if not isinstance(w, np.ndarray):
            continue
        g = host_of(g)
        if w.dtype.kind in "iu":
            assert np.array_equal(w, g), f"{name}[{real}] integer arg {i} differs"
            continue
        scale = max(float(np.nanmax(np.abs(w))), 1e-30)
        err = float(np.nanmax(np.abs(g.astype(np.float64) - w.astype(np.float64)))) / scale
        assert np.array_equal(np.isnan(w), np.isnan(g)), f"{name}[{real}] arg {i}: NaN pattern"
        assert err <= TOL[real], f"{name}[{real}] arg {i}: rel err {err:.3e}"
    if want_ret is not None:
        w = np.atleast_1d(np.asarray(want_ret, dtype=np.float64))
        g = np.atleast_1d(np.asarray(got_ret, dtype=np.float64))
        if name in ("lsm3ddeterminenarrowband_", "lsm3dcheckouternarrowbandlayer_"):
            assert np.array_equal(w, g), f"{name}[{real}] integer outputs {g} != {w}"
        else:
            t = SUM_TOL[real] if "avegrad" in name else TOL[real]
            err = float(np.max(np.abs(g - w) / np.maximum(np.abs(w), 1e-30)))
            assert err <= t, f"{name}[{real}] scalar: got {g} want {w} (rel {err:.3e})"


@pytest.mark.gpu
@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("name", sorted(LOCAL))
def test_local_device_matches_oracle(name, real, device):
    import torch
    lib, ora = toolbox(real), oracle_lib(real)
    for case in local_cases_for(name, real):
        want = case.copy()
        rw = ora.call(name, *want.args)
        dev = to_device(name, case.args, device)
        rg = lib.call(name, *dev)
        torch.cuda.synchronize()
        compare(name, real, want.args, rw, dev, rg)


@pytest.mark.gpu
@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("name", sorted(LOCAL))
def test_local_host_memory_matches_device_memory(name, real, device):
    """Host arrays (staged by the library) give the device results bit for bit."""
    import torch
    lib = toolbox(real)
    case = local_cases_for(name, real, seeds=(1,))[0]
    host = case.copy()
    rh = lib.call(name, *host.args)
    dev = to_device(name, case.args, device)
    rd = lib.call(name, *dev)
    torch.cuda.synchronize()
    for i, (h, d) in enumerate(zip(host.args, dev)):
        if isinstance(h, np.ndarray):
            assert np.array_equal(h, host_of(d), equal_nan=True), f"{name}[{real}] arg {i}"
    if rh is not None:
        assert np.array_equal(np.atleast_1d(rh), np.atleast_1d(rd))


def _band_args(real, n_side, seed, device):
    import torch
    box = (0, n_side - 1, -2, n_side - 3, 1, n_side)
    dx = 1.0 / n_side
    rng = np.random.default_rng(seed)
    shape = box_shape(box)
    k, j, i = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
    r = np.sqrt(((i - 0.47 * n_side) * dx) ** 2 + ((j - 0.52 * n_side) * dx) ** 2 + ((k - 0.5 * n_side) * dx) ** 2)
    phi = (r - 0.31 + 2e-3 * rng.standard_normal(shape)).astype(np.float64 if real == "double" else np.float32)
    n = phi.size
    host = [phi, box, np.zeros(shape, np.uint8), box, *[np.full(n, -7, np.int32) for _ in range(3)], 0, n - 1,
            np.full(6, -9, np.int32), np.full(6, -9, np.int32), np.full(n, -7, np.int32), 0, n - 1, 4 * dx, 2 * dx,
            LEVEL]
    return host, box


@pytest.mark.gpu
@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("n_side", [48, 160])
def test_band_builder_at_size_matches_oracle(real, n_side, device):
    """Many blocks of the ordered compaction and of the neighbour bids."""
    import torch
    lib, ora = toolbox(real), oracle_lib(real)
    host, box = _band_args(real, n_side, 3, device)
    want = [a.copy() if isinstance(a, np.ndarray) else a for a in host]
    rw = ora.lsm3ddeterminenarrowband_(*want)
    dev = to_device("lsm3ddeterminenarrowband_", host, device)
    rg = lib.lsm3ddeterminenarrowband_(*dev)
    torch.cuda.synchronize()
    assert tuple(rw) == tuple(rg)
    for w, g in zip(want, dev):
        if isinstance(w, np.ndarray) and w.dtype.kind in "iu":
            assert np.array_equal(w, host_of(g))
    n_band = int(want[10][LEVEL] + 1)
    assert n_band > 0.05 * want[0].size


@pytest.mark.gpu
def test_band_levels_from_mask_at_size(device):
    """DETERMINE_NARROW_BAND_FROM_MASK over a large box, both mask signs."""
    import torch
    lib, ora = toolbox("double"), oracle_lib("double")
    host, box = _band_args("double", 96, 5, device)
    for sign in (1, -1):
        n = host[0].size
        args = [host[0], box, np.zeros(box_shape(box), np.uint8), box, *[np.full(n, -7, np.int32) for _ in range(3)],
                0, n - 1, np.full(6, -9, np.int32), np.full(6, -9, np.int32), 2, sign]
        want = [a.copy() if isinstance(a, np.ndarray) else a for a in args]
        ora.lsm3ddeterminenarrowbandfrommask_(*want)
        dev = to_device("lsm3ddeterminenarrowbandfrommask_", args, device)
        lib.lsm3ddeterminenarrowbandfrommask_(*dev)
        torch.cuda.synchronize()
        for w, g in zip(want, dev):
            if isinstance(w, np.ndarray) and w.dtype.kind in "iu":
                assert np.array_equal(w, host_of(g))


@pytest.mark.gpu
def test_entries_outside_the_arrays_are_reported(device):
    """A list entry outside phi's box is skipped (no access) and reported."""
    import torch
    lib = toolbox("double")
    lib.set_sync(True)   # reported when a call completes synchronously (drivers may have left it off)
    fb = grow(GB, -3)
    rhs = torch.zeros(box_shape(GB), dtype=torch.float64, device=device)
    ix = torch.tensor([0, 1, GB[1] + 4], dtype=torch.int32, device=device)
    iy = torch.tensor([0, 1, 0], dtype=torch.int32, device=device)
    iz = torch.tensor([3, 3, 3], dtype=torch.int32, device=device)
    lib.lsm3dzerooutlevelseteqnrhslocal_(rhs, GB, ix, iy, iz, 0, 1)   # in range: fine
    with pytest.raises(LsmError, match="outside"):
        lib.lsm3dzerooutlevelseteqnrhslocal_(rhs, GB, ix, iy, iz, 0, 2)
    # the counter was cleared: a good call after the failure succeeds
    lib.lsm3dzerooutlevelseteqnrhslocal_(rhs, GB, ix, iy, iz, 0, 1)
    assert fb is not None


@pytest.mark.gpu
def test_index_arrays_too_short_is_an_error(device):
    import torch
    lib = toolbox("double")
    rng = np.random.default_rng(0)
    phi = torch.from_numpy(sphere("double", rng)).to(device)
    nb = torch.zeros(box_shape(GB), dtype=torch.uint8, device=device)
    short = [torch.full((10,), -7, dtype=torch.int32, device=device) for _ in range(3)]
    n_lo, n_hi = np.full(6, -9, np.int32), np.full(6, -9, np.int32)
    outer = torch.full((4000,), -7, dtype=torch.int32, device=device)
    with pytest.raises(LsmError, match="too short"):
        lib.lsm3ddeterminenarrowband_(phi, GB, nb, GB, *short, 0, 9, n_lo, n_hi, outer, 0, 3999, GAMMA, BETA, LEVEL)


@pytest.mark.gpu
def test_band_then_local_step_matches_oracle(device):
    """A band built on the GPU feeds the local kernels (the local example's
    sequence: marks, ENO2 on levels 0-2, the normal-speed term, RK1)."""
    import torch
    real = "double"
    lib, ora = toolbox(real), oracle_lib(real)
    rng = np.random.default_rng(11)
    phi = sphere(real, rng)
    n = phi.size

    def run(L, arr):
        nb = arr(np.zeros(box_shape(GB), np.uint8))
        lists = [arr(np.full(n, -7, np.int32)) for _ in range(3)]
        n_lo, n_hi = np.full(6, -9, np.int32), np.full(6, -9, np.int32)
        outer = arr(np.full(n, -7, np.int32))
        p = arr(phi.copy())
        L.lsm3ddeterminenarrowband_(p, GB, nb, GB, *lists, 0, n - 1, n_lo, n_hi, outer, 0, n - 1, GAMMA, BETA, LEVEL)
        for box, mark in ((grow(GB, -2), MARK_D2), (grow(GB, -1), MARK_D1), (GB, MARK_GB)):
            L.lsm3dmarknarrowbandboundarylayer_(nb, GB, box, mark)
        g = [arr(np.zeros_like(phi)) for _ in range(8)]
        L.lsm3dhjeno2local_(*g[:3], GB, *g[3:6], GB, p, GB, g[6], GB, g[7], GB, 0.1, 0.13, 0.07, *lists,
                            int(n_lo[0]), int(n_hi[0]), int(n_lo[1]), int(n_hi[1]), int(n_lo[2]), int(n_hi[2]), nb, GB,
                            MARK_FB, MARK_D1, MARK_D2)
        rhs = arr(np.zeros_like(phi))
        L.lsm3dzerooutlevelseteqnrhslocal_(rhs, GB, *lists, int(n_lo[0]), int(n_hi[0]))
        L.lsm3daddconstnormalveltermtolserhslocal_(rhs, GB, *g[:3], GB, *g[3:6], GB, 0.4, *lists, int(n_lo[0]),
                                                   int(n_hi[0]), nb, GB, MARK_FB)
        out = arr(phi.copy())
        L.lsm3drk1steplocal_(out, GB, p, GB, rhs, GB, 0.01, *lists, int(n_lo[0]), int(n_hi[0]), nb, GB, MARK_FB)
        return [host_of(a) for a in (out, rhs, *g)], (n_lo.copy(), n_hi.copy())

    want, lw = run(ora, lambda a: a)
    got, lg = run(lib, lambda a: torch.from_numpy(a).to(device))
    torch.cuda.synchronize()
    assert all(np.array_equal(a, b) for a, b in zip(lw, lg))
    for w, g in zip(want, got):
        assert np.array_equal(w, g)
