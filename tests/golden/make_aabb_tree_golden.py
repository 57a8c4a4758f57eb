"""Regenerate tests/golden/aabb_tree_golden.npz: the reference's own bounding
box tree (boxes, nodes, face order, stack size) on the meshes of its AABB
tests.

The tree layout is not fixed by any closed form or brute force (only the
distances are), so its expected arrays come from the reference builder
itself: `_build_aabb` of python/igl/__init__.py:286-319, plain NumPy,
imported from the reference checkout under another package name (its module
imports only numpy, scipy and its ctypes shim, which loads no library at
import). Meshes: grid(8), grid(9), grid(10), grid(60) of
tests/test_parity.py:12-26 (the AABB tests' meshes, :131-177), a flat grid
whose box centres tie on two axes, and a seeded triangle soup; leaf sizes 8
(the reference default) and 3.

    python tests/golden/make_aabb_tree_golden.py [/root/reference]
"""

from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "mojo-libigl_amd", "python"))

import wn_workloads as W  # noqa: E402


def reference_builder(checkout: str):
    pkg = os.path.join(checkout, "python", "igl")
    spec = importlib.util.spec_from_file_location(
        "reference_igl", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    module = importlib.util.module_from_spec(spec)
    sys.modules["reference_igl"] = module
    spec.loader.exec_module(module)
    return module._build_aabb


def meshes():
    out = {f"grid{n}": W.grid(n) for n in (8, 9, 10, 60)}
    out["flat12"] = W.grid(12, curved=False)
    rng = np.random.default_rng(23)
    V = rng.normal(size=(200, 3))
    out["soup"] = (V, rng.integers(0, 200, size=(700, 3)))
    return out


def main() -> None:
    checkout = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    build = reference_builder(checkout)
    out = {}
    for name, (V, F) in meshes().items():
        V = np.ascontiguousarray(V, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.int64)
        out[f"{name}_V"], out[f"{name}_F"] = V, F
        for leaf in (8, 3):
            boxes, nodes, order, stack = build(V, F, leaf)
            key = f"{name}_leaf{leaf}"
            out[f"{key}_boxes"], out[f"{key}_nodes"] = boxes, nodes
            out[f"{key}_order"], out[f"{key}_stack"] = order, np.int64(stack)
    path = os.path.join(HERE, "aabb_tree_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main()
