"""Linear-system capture files (.cfdxlin, CFDX_LINEAR_REPLAY_V1).

Pinned against the reference's own writer and reader (oracle/_ref, built
from core/linalg/linear_system_capture.h): files are byte-identical, each
side reads the other's; the checksum is also recomputed here independently
(64-bit FNV-1a over little-endian 64-bit words, in the documented order).
The replay case solves a captured pressure system on the GPU and compares
with the oracle's PCG.
"""
import struct

import numpy as np
import pytest

import cfdx_dev_amd as cx
from systems import box_mesh, pressure_system


def fnv1a_words(words):
    h = 14695981039346656037
    for w in words:
        for byte in struct.pack("<Q", w & 0xFFFFFFFFFFFFFFFF):
            h ^= byte
            h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def _cap(oracle, dims=(6, 5, 4), seed=3):
    s = pressure_system(oracle, box_mesh(*dims, seed=seed), sigma=1.0, seed=seed)
    x0 = np.random.default_rng(seed).normal(size=s["n"]) * 1e-3
    return cx.LinearSystemCapture(s["rowp"].astype(np.uint32), s["cols"].astype(np.uint32), s["vals"], s["b"], x0,
                                  problem=1, null_space=0, scaling=0, max_iterations=500, tolerance=1e-10)


def test_round_trip_and_checksum(oracle, tmp_path):
    c = _cap(oracle)
    p = tmp_path / "a.cfdxlin"
    chk = cx.write_capture(p, c)
    r = cx.read_capture(p)
    for f in ("rowp", "cols", "vals", "rhs", "initial_solution"):
        assert np.array_equal(getattr(r, f), getattr(c, f)), f
    assert (r.problem, r.null_space, r.scaling, r.max_iterations, r.tolerance) == (1, 0, 0, 500, 1e-10)
    words = [c.n, c.n, len(c.vals), 1, 0, 0, 500, bits(1e-10)] + [int(v) for v in c.rowp]
    for k in range(len(c.vals)):
        words += [int(c.cols[k]), bits(c.vals[k])]
    for i in range(c.n):
        words += [bits(c.rhs[i]), bits(c.initial_solution[i])]
    assert chk == r.checksum == fnv1a_words(words)
    text = p.read_text().split("\n")
    assert text[0] == "CFDX_LINEAR_REPLAY_V1" and text[1] == "1 0 0"
    assert text[2] == f"{c.n} {c.n} {len(c.vals)} 500" and text[3] == f"{bits(1e-10):x}"
    assert text[-2] == f"{chk:x}" and text[-1] == ""


def test_byte_identical_to_reference(oracle, reference, tmp_path):
    from oracle import ref_capture_read, ref_capture_write

    c = _cap(oracle, (7, 6, 5), 4)
    ours, theirs = tmp_path / "ours.cfdxlin", tmp_path / "ref.cfdxlin"
    cx.write_capture(ours, c)
    assert ref_capture_write(theirs, c.rowp, c.cols, c.vals, c.rhs, c.initial_solution, 1, 0, 0, 500, 1e-10)
    assert ours.read_bytes() == theirs.read_bytes()
    back = ref_capture_read(ours)
    assert back is not None and back["checksum"] == cx.read_capture(theirs).checksum
    for a, b in (("rowp", "rowp"), ("cols", "cols"), ("vals", "vals"), ("rhs", "rhs"), ("x0", "initial_solution")):
        assert np.array_equal(back[a], getattr(c, b))


def test_rejections(oracle, tmp_path):
    c = _cap(oracle)
    good = tmp_path / "g.cfdxlin"
    cx.write_capture(good, c)
    lines = good.read_text().split("\n")

    def bad(name, text, reason):
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(ValueError, match=reason):
            cx.read_capture(p)

    k = 4 + c.n + 1 + 3  # a nonzero's line
    col, val = lines[k].split()
    flipped = lines[:k] + [f"{col} {int(val, 16) ^ 1:x}"] + lines[k + 1:]
    bad("flip.cfdxlin", "\n".join(flipped), "checksum mismatch")
    bad("trunc.cfdxlin", "\n".join(lines[: len(lines) // 2]), "truncated")
    bad("trail.cfdxlin", "\n".join(lines) + "ff\n", "trailing")
    bad("tag.cfdxlin", "\n".join(["CFDX_LINEAR_REPLAY_V2"] + lines[1:]), "unsupported format")
    bad("meta.cfdxlin", "\n".join(lines[:1] + ["6 0 0"] + lines[2:]), "metadata")
    bad("big.cfdxlin", "\n".join(lines[:2] + [f"2000001 2000001 {len(c.vals)} 500"] + lines[3:]), "limits")
    unsorted = cx.LinearSystemCapture(**{**c.__dict__})
    unsorted.cols = c.cols.copy()
    unsorted.cols[[0, 1]] = unsorted.cols[[1, 0]]
    with pytest.raises(ValueError, match="ascending"):
        cx.write_capture(tmp_path / "u.cfdxlin", unsorted)
    zero_tol = cx.LinearSystemCapture(**{**c.__dict__})
    zero_tol.tolerance = 0.0
    with pytest.raises(ValueError, match="iteration controls"):
        cx.write_capture(tmp_path / "t.cfdxlin", zero_tol)


def test_reference_rejects_what_we_reject(oracle, reference, tmp_path):
    from oracle import ref_capture_read

    c = _cap(oracle)
    good = tmp_path / "g.cfdxlin"
    cx.write_capture(good, c)
    lines = good.read_text().split("\n")
    k = 4 + c.n + 1 + 3
    col, val = lines[k].split()
    p = tmp_path / "flip.cfdxlin"
    p.write_text("\n".join(lines[:k] + [f"{col} {int(val, 16) ^ 1:x}"] + lines[k + 1:]))
    assert ref_capture_read(p) is None


@pytest.mark.gpu
def test_replay_on_gpu_matches_oracle(oracle, tmp_path):
    """A captured pressure system replayed through the GPU AMG-PCG."""
    c = _cap(oracle, (14, 12, 10), 6)
    p = tmp_path / "replay.cfdxlin"
    cx.write_capture(p, c)
    r = cx.read_capture(p)
    a = oracle.amg("sa")
    assert a.setup(r.n, r.rowp, r.cols, r.vals)
    xo = r.initial_solution.copy()
    assert a.pcg(r.rhs, xo, r.max_iterations, r.tolerance).status == 0
    h = cx.NativeAMGPreconditioner()
    assert h.setup(r.n, r.rowp, r.cols, r.vals), h.last_error()
    xg = r.initial_solution.copy()
    res = h.solve_cg(r.rhs, xg, r.max_iterations, r.tolerance)
    assert res.status == 0
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.gpu
def test_replay_tool(oracle, tmp_path):
    from cfdx_dev_amd.replay import replay

    c = _cap(oracle, (10, 9, 8), 7)
    p = tmp_path / "tool.cfdxlin"
    cx.write_capture(p, c)
    for solver in ("cg", "fgmres", "mixed"):
        out = replay(p, solver=solver)
        assert out["status"] == 0 and out["residual_relative"] <= c.tolerance, (solver, out)
