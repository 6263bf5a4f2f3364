"""Writes the committed golden fixtures of tests/golden/ from the CPU oracle.

Run from the repo root after building oracle/liboracle.so:
    python tests/golden/make_golden.py
The fixtures are regression anchors: SHA-256 digests of the synthetic
inputs and of the oracle's outputs for the 352x288 five-frame group
(strength 3, centre 2) and for the sub-pel MC / SAD workload on it.
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "oxideav-vp8_amd"), str(ROOT / "oracle")]

import vp8g.synth as synth  # noqa: E402
from oracle_ctypes import LumaRef, Oracle, RefPlanes  # noqa: E402


def sha(*arrs) -> str:
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main() -> None:
    orc = Oracle()
    g = synth.arnr_group_cif()
    inp = sha(*[np.concatenate([y.reshape(-1), u.reshape(-1), v.reshape(-1)])
                for y, u, v in zip(g.y, g.u, g.v)])
    st, out, _, mvs = orc.build_arnr_altref(g.frames(), g.center, 3, threads=4, want_mvs=True)
    assert st == 0
    arnr = {"workload": "arnr_group_cif 352x288 n=5 center=2 strength=3",
            "input_sha256": inp, "output_sha256": sha(*out), "mvs_sha256": sha(mvs)}
    (ROOT / "tests/golden/arnr_cif.json").write_text(json.dumps(arnr, indent=1) + "\n")

    mc = synth.mc_workload(g, config=0)
    rp = RefPlanes(mc.ref_y.ctypes.data, mc.ref_u.ctypes.data, mc.ref_v.ctypes.data,
                   mc.mb_cols * 16, mc.mb_cols * 8, mc.mb_cols, mc.mb_rows)
    n = len(mc.mb_xy)
    res = {"workload": "mc_workload(arnr_group_cif) 22x18 MBs, mv in [-64,64] quarter-pel"}
    for name, filt in (("sixtap", orc.SIXTAP), ("bilinear", orc.BILINEAR)):
        for full in (0, 1):
            out = np.zeros((n, 384), np.uint8)
            f = np.ascontiguousarray(filt, np.int32)
            orc.lib.orc_predict_inter_mb_batch(rp, mc.mb_xy.ctypes.data, mc.mvs.ctypes.data, n, full,
                                               f.ctypes.data, out.ctypes.data)
            res[f"predict_{name}_full{full}_sha256"] = sha(out)
    lr = LumaRef(mc.ref_y.ctypes.data, mc.mb_cols * 16, mc.mb_cols * 16, mc.mb_rows * 16)
    sad = np.zeros(n, np.uint32)
    orc.lib.orc_mb_luma_sad_at_mv_batch(lr, mc.mb_xy.ctypes.data, mc.src.ctypes.data,
                                        mc.mvs.ctypes.data, n, sad.ctypes.data)
    res["sad_sha256"] = sha(sad)
    res["sad_sum"] = int(sad.astype(np.int64).sum())
    (ROOT / "tests/golden/mc_cif.json").write_text(json.dumps(res, indent=1) + "\n")
    print("wrote tests/golden/arnr_cif.json, tests/golden/mc_cif.json")


if __name__ == "__main__":
    main()
