"""Invariances of the exact winding number on every GPU path, each checked
against the oracle evaluated on the same transformed inputs:

* translation far from the origin (1e4 .. 1e6): the FP32-compensated paths
  keep their bar only because each block works in FP32 offsets from its own
  origin; plain FP32 coordinates at 1e6 would be off by ~0.06;
* uniform scale (1e-5, 1e-3, 1e3, 5e3, 1e4, 2e4, 1e6: unscaled, the FP32
  product of a group of faces' den would underflow or overflow at the ends,
  or land between 2^126 and the FP32 maximum, where its reciprocal is no
  longer a normal number; the kernels stage every tile in a power-of-two
  frame that keeps the products in range), a rotation, a face permutation,
  a cyclic permutation of every face's corners (w unchanged), and a swap of
  two corners of every face (w negated);
* a power-of-two scale (2^-40 .. 2^40) of the whole input leaves the
  FP32-compensated results bit for bit unchanged on a closed mesh, and
  within FP64 rounding (1e-15) on an open one, whose boundary correction
  adds FP64 closed forms through the library atan2.

Paths: FP64 face kernel (wn_winding_gpu_vf, the path of igl.winding_number),
FP32c face kernel, FP32c strip kernel (igl.Mesh), FP64 strip kernel
(igl.Mesh, precision="fp64"). Mesh: C1 (closed, 5,120 faces), 1,024
queries."""

import numpy as np
import pytest

import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

BARS = {"fp64_face": 1e-10, "fp32c_face": 1e-6, "fp32c_strip": 1e-6, "fp64_strip": 1e-10}


def run(path, V, F, Q):
    import igl
    from igl._lib import addr

    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int64)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    if path in ("fp64_face", "fp32c_face"):
        out = np.full(len(Q), np.nan)
        status = igl.lib().wn_winding_gpu_vf(addr(V), addr(F), len(V), len(F), 0, addr(Q),
                                             addr(out), len(Q), 0 if path == "fp64_face" else 1)
        assert status == 0, igl._lib.last_error()
        return out
    m = igl.Mesh(V, F)
    try:
        return m.winding_number(Q, precision="fp64" if path == "fp64_strip" else "fp32c")
    finally:
        m.close()


def rotation(seed):
    q = np.random.default_rng(seed).normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def transforms():
    V, F, Q = W.config1()
    perm = np.random.default_rng(7).permutation(len(F))
    R = rotation(11)
    return {
        "translate_1e4": (V + [1e4, -2e4, 3e4], F, Q + [1e4, -2e4, 3e4], 1.0),
        "translate_1e6": (V + 1e6, F, Q + 1e6, 1.0),
        "scale_1e-3": (V * 1e-3, F, Q * 1e-3, 1.0),
        "scale_1e3": (V * 1e3, F, Q * 1e3, 1.0),
        # the FP32c fast paths multiply three faces' den (~ size^3 each):
        # the product underflows at 1e-5 and overflows at 1e6, and those
        # groups must fall back to one face at a time
        "scale_1e-5": (V * 1e-5, F, Q * 1e-5, 1.0),
        "scale_1e6": (V * 1e6, F, Q * 1e6, 1.0),
        # products of 2^126 .. 3.4e38: finite, but 1 / re is below the
        # normal range (the reciprocal unit returns 0 there)
        "scale_5e3": (V * 5e3, F, Q * 5e3, 1.0),
        "scale_1e4": (V * 1e4, F, Q * 1e4, 1.0),
        "scale_2e4": (V * 2e4, F, Q * 2e4, 1.0),
        "rotate": (V @ R.T, F, Q @ R.T, 1.0),
        "permute_faces": (V, F[perm], Q, 1.0),
        "cycle_corners": (V, F[:, [1, 2, 0]], Q, 1.0),
        "swap_corners": (V, F[:, [0, 2, 1]], Q, -1.0),
    }


@pytest.fixture(scope="module")
def base():
    V, F, Q = W.config1()
    return oracle.winding(V, F, Q)


@pytest.mark.parametrize("path", list(BARS))
@pytest.mark.parametrize("name", list(transforms()))
def test_invariance(require_gpu, base, path, name):
    V, F, Q, sign = transforms()[name]
    ref = oracle.winding(V, F, Q)
    w = run(path, V, F, Q)
    err = np.abs(w - ref).max()
    # the oracle itself on the transformed inputs: the transform keeps w up
    # to the rounding of the transformed coordinates
    drift = np.abs(ref - sign * base).max()
    print(f"{path} {name}: max|w - oracle| {err:.3e}, oracle drift {drift:.3e}", flush=True)
    assert err <= BARS[path]
    assert drift <= 1e-12


@pytest.mark.parametrize("path", ["fp32c_face", "fp32c_strip"])
@pytest.mark.parametrize("mesh", ["closed", "open"])
def test_power_of_two_scale_is_exact(require_gpu, path, mesh):
    """Every FP32 quantity of the FP32-compensated kernels is computed in a
    power-of-two frame of its tile, and FP32 rounding commutes with a
    power-of-two factor, so scaling V and Q by 2^k leaves every angle, and w,
    unchanged bit for bit (for coordinates that stay normal numbers). The
    open mesh adds the boundary correction, whose FP64 closed forms call the
    library atan2: equal there to FP64 rounding."""
    if mesh == "closed":
        V, F, Q = W.config1()
    else:
        V, F = W.bench_grid(40)
        Q = W.padded_box_queries(V, 1024, np.random.default_rng(5))
    w0 = run(path, V, F, Q)
    ref = oracle.winding(V, F, Q)
    assert np.abs(w0 - ref).max() <= BARS[path]
    for k in (-40, -12, 13, 40):
        s = 2.0 ** k
        w = run(path, V * s, F, Q * s)
        diff = np.abs(w - w0).max()
        print(f"{path} {mesh} 2^{k}: max|w - w(2^0)| {diff:.3e}", flush=True)
        if mesh == "closed":
            assert np.array_equal(w, w0), (k, diff)
        else:
            assert diff <= 1e-15, (k, diff)


def soup(seed):
    """Non-manifold triangle soup: edges shared by 3+ faces, duplicate and
    reversed duplicate faces, faces with a repeated vertex, unused vertices."""
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(400, 3))
    F = rng.integers(0, 300, size=(6000, 3))
    F[::23, 2] = F[::23, 0]                            # repeated vertex
    F = np.vstack([F, F[:200], F[200:400, ::-1]])      # duplicates, reversed duplicates
    Q = rng.uniform(-3, 3, size=(2048, 3))
    return V, F, Q


@pytest.mark.parametrize("path", list(BARS))
@pytest.mark.parametrize("seed", [1, 2])
def test_triangle_soup(require_gpu, path, seed):
    """Every GPU path on a non-manifold soup against the oracle: FP64 at its
    bar, FP32c at 2e-6. The soup's 6,400 random faces overlap at every query
    (sum of |angle| over the faces ~440 rad per query, against 2 pi for a
    closed surface), so the FP32 rounding of each pair's angle (~eps |angle|)
    adds up to ~3.4e-7 rms in w; the boundary rounding itself is corrected
    (before the boundary correction these soups read up to 7.3e-6)."""
    V, F, Q = soup(seed)
    ref = oracle.winding(V, F, Q)
    err = np.abs(run(path, V, F, Q) - ref).max()
    bar = BARS[path] if path.startswith("fp64") else 2e-6
    print(f"soup {seed} {path}: max|w - oracle| {err:.3e}", flush=True)
    assert err <= bar
