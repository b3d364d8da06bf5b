"""Comparison of the HIP result against the oracle's (test helper)."""
import numpy as np


def first_diff(a, b):
    idx = np.flatnonzero(a != b)
    return None if idx.size == 0 else int(idx[0])


def compare_results(res, ref, levels=None, ref_mips=None, gpu_levels=None):
    """Bit-exact comparison.  Colour (RGBA8) and normal (RGB8) bytes are
    compared exactly; the north-star tolerance is one quantisation step, so a
    mismatch larger than one step is reported separately from one that is
    within it."""
    errors = []
    assert tuple(res["S"]) == tuple(ref.S), (res["S"], ref.S)
    assert res["levels"] == ref.levels
    if res["occupancy"] is not None:
        d = first_diff(res["occupancy"], ref.occupancy)
        if d is not None:
            errors.append(f"occupancy word {d}: {res['occupancy'][d]:#x} vs {ref.occupancy[d]:#x}")
    for dd in range(ref.levels):
        g, r = res["tree"][dd], ref.tree[dd]
        for key in ("types", "const", "bidx"):
            d = first_diff(g[key], r[key])
            if d is not None:
                errors.append(f"depth {dd} {key}[{d}]: {g[key][d]} vs {r[key][d]}")
        if g["nb"] != r["nb"]:
            errors.append(f"depth {dd} bricks {g['nb']} vs {r['nb']}")
            continue
        for key in ("bcol", "bnrm"):
            if key not in g:
                continue
            a = g[key].view(np.uint8).astype(np.int16)
            b = r[key].view(np.uint8).astype(np.int16)
            if a.shape != b.shape:
                errors.append(f"depth {dd} {key} shape {a.shape} vs {b.shape}")
                continue
            diff = np.abs(a - b)
            if diff.max(initial=0) > 1:
                errors.append(f"depth {dd} {key} differs by {int(diff.max())} steps "
                              f"({int((diff > 1).sum())} bytes)")
            elif diff.max(initial=0) == 1:
                errors.append(f"depth {dd} {key} within one step but not bit-exact "
                              f"({int((diff > 0).sum())} bytes)")
    if gpu_levels is not None and ref_mips:
        for k, (gc, gn) in enumerate(gpu_levels):
            rc, rn = ref_mips[k]
            d = first_diff(gc.ravel(), rc.ravel())
            if d is not None:
                errors.append(f"level {k} colour voxel {d}: {gc.ravel()[d]:#x} vs {rc.ravel()[d]:#x}")
            gb, rb = gn.view(np.uint32).ravel(), rn.view(np.uint32).ravel()
            d = first_diff(gb, rb)
            if d is not None:
                errors.append(f"level {k} normal component {d}: {gn.ravel()[d]!r} vs {rn.ravel()[d]!r}")
    return errors
