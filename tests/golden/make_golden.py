"""Write the golden fixtures from the reference build (oracle/_ref/mts_ref).

Run in the build container, where /root/reference exists and
`make -C oracle ref` has built the reference solver from its own sources:

    python tests/golden/make_golden.py

Each case becomes tests/golden/<name>.npz holding
  meta         JSON of the case (method, mib, n, dt, T, surface, beta)
  cr{d}_int    per crossing of direction d (0 x, 1 y, 2 z), in the product's
               line order: j, k, ID, left_loc, nw
  cr{d}_dbl    gamma, coord(3), normal(3), beta_in, beta_out, P(9), weil(9), weir(9)
  u_steps      fields after steps 1..K, float64 (K, nz, ny, nx), K = 3 (2 on the
               33^3 grids), for per-step checks restarted from the reference's
               own field (the tests read u_steps[0] as u_step1)
  u_final      field after the last step
  summary      steps, t_final, lmax, l2, argmax(3)
  jumps        [u] and u_dir per crossing after step 1's Refresh_Jump, in the
               reference's dump order (z lines, y lines, x lines)
  diag{d}      per crossing of direction d, in the product's line order: the
               relative errors of its two fictitious points and of its jump
               u_dir at the final time (Refresh_Fp, Refresh_Jump)
  diag_report  the reference's Error_Fp and Error_Jump lines (FP Lmax, L2,
               jump Lmax, L2; printed with 3 digits)
The vectors are data produced by the reference binary; no reference source is
stored here.
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mts_oracle as orc  # noqa: E402

R0 = 0.5 * (1 - 2 ** -20)  # sphere radius pulled off the grid node (0.5, 0, 0)
E1 = 1 - 2 ** -20          # ellipsoid scale pulled off the grid node (0, 0.5, 0)

CASES = []
for _m in ("adi", "lod-ie", "lod-cn", "ts"):
    for _mib in (1, 2):
        CASES.append(dict(name=f"sphere16_{_m}_mib{_mib}", method=_m, mib=_mib, n=16, dt=0.125,
                          T=0.5, surface=("ellipsoid", 1, 1, 1, R0), beta=("const", 10, 1)))
CASES += [
    # the 33^3 Douglas-ADI workload: MIB-L1, sphere r=0.5 (pulled off the grid), 16 steps
    dict(name="cfg0_adi_mib1", method="adi", mib=1, n=32, dt=0.0625, T=1.0,
         surface=("ellipsoid", 1, 1, 1, R0), beta=("const", 10, 1)),
    # the 257^3 LOD-CN workload's geometry and physics at 33^3 (the reference cannot set
    # up 257^3: its tangential-jump diagnostics exit for near-tangent crossings)
    dict(name="cfg1like32_lodcn_mib2", method="lod-cn", mib=2, n=32, dt=1 / 16, T=0.5,
         surface=("ellipsoid", 0.7, 0.5, 0.4, E1), beta=("sincos", 10, 1)),
    # the 513^3 TS workload's physics (beta 1000/1, bumpy sphere) at 27^3
    dict(name="bumpy26_ts_mib2", method="ts", mib=2, n=26, dt=1 / 13, T=0.25,
         surface=("bumpy", 0.5 * (1 - 2 ** -20), 0.1, 5, 3), beta=("const", 1000, 1)),
    # the batch and serving mixes' geometry family: torus, strong contrast either way
    dict(name="torus22_lodcn_mib2", method="lod-cn", mib=2, n=22, dt=1 / 11, T=0.5,
         surface=("torus", 0.5, 0.2), beta=("const", 0.01, 1)),
    dict(name="torus22_lodie_mib1", method="lod-ie", mib=1, n=22, dt=1 / 11, T=0.5,
         surface=("torus", 0.5, 0.17), beta=("const", 100, 1)),
    dict(name="ellipsoid20_adi_mib2", method="adi", mib=2, n=20, dt=0.1, T=0.5,
         surface=("ellipsoid", 0.8, 0.6, 0.5, E1), beta=("const", 100, 1)),
    # Douglas-ADI on interfaces with corner pairs (32 corner crossings along z on
    # the torus, 8 along y and 8 along z on the 257^3 LOD-CN workload's ellipsoid)
    dict(name="torus22_adi_mib1", method="adi", mib=1, n=22, dt=1 / 11, T=0.5,
         surface=("torus", 0.5, 0.17), beta=("const", 100, 1)),
    dict(name="torus22_adi_mib2", method="adi", mib=2, n=22, dt=1 / 11, T=0.5,
         surface=("torus", 0.5, 0.17), beta=("const", 0.01, 1)),
    dict(name="cfg1like32_adi_mib1", method="adi", mib=1, n=32, dt=1 / 16, T=0.5,
         surface=("ellipsoid", 0.7, 0.5, 0.4, E1), beta=("sincos", 10, 1)),
    dict(name="cfg1like32_adi_mib2", method="adi", mib=2, n=32, dt=1 / 16, T=0.5,
         surface=("ellipsoid", 0.7, 0.5, 0.4, E1), beta=("sincos", 10, 1)),
]


def steps_kept(case):
    """Per-step fields kept in the fixture (bounded by the fixture size)."""
    return 2 if case["n"] >= 32 else 3


def main():
    if not orc.ref_available():
        raise SystemExit("oracle/_ref/mts_ref is missing: run make -C oracle ref")
    tmp = tempfile.TemporaryDirectory(prefix="mts_golden_")
    work = tmp.name
    for case in CASES:
        s = case["surface"]
        b = case["beta"]
        sstr = f"{s[0]}:" + ",".join(repr(float(v)) for v in s[1:])
        bstr = f"{b[0]}:" + ",".join(repr(float(v)) for v in b[1:])
        out = os.path.join(work, case["name"])
        try:
            ref = orc.run_reference(out, case["method"], case["mib"], case["n"], case["dt"],
                                    case["T"], sstr, bstr, dump_steps=steps_kept(case))
        except RuntimeError as e:
            print(f"{case['name']}: skipped ({e})")
            continue
        nn = case["n"] + 1
        arrays = {"meta": np.array(json.dumps(case))}
        diag_path = os.path.join(out, "diag.txt")
        diag = np.loadtxt(diag_path, ndmin=2) if os.path.getsize(diag_path) else np.zeros((0, 3))
        # the dump runs over z lines, then y lines, then x lines
        counts = [len(ref["crossings"][d]) for d in range(3)]
        starts = {2: 0, 1: counts[2], 0: counts[2] + counts[1]}
        for d in range(3):
            cr0 = ref["crossings"][d]
            dd0 = diag[starts[d]:starts[d] + counts[d]]
            if len(cr0):
                order = np.argsort(cr0["k"].astype(np.int64) * nn + cr0["j"], kind="stable")
                dd0 = dd0[order]
            arrays[f"diag{d}"] = dd0.astype(np.float64)
        rep = open(os.path.join(out, "diag_report.txt")).read().split()
        nums = [float(rep[i + 2]) for i, w in enumerate(rep) if w in ("Lmax", "L2") and rep[i + 1] == "="]
        arrays["diag_report"] = np.array(nums, dtype=np.float64)
        for d in range(3):
            nj = nn  # all axes have n + 1 nodes
            cr = orc.sort_by_line(ref["crossings"][d], nj)
            ii = np.stack([cr["j"], cr["k"], cr["id"], cr["left_loc"], cr["nw"]], axis=1) \
                if len(cr) else np.zeros((0, 5), np.int32)
            dd = np.concatenate([cr["gamma"][:, None], cr["coord"], cr["normal"],
                                 cr["beta_in"][:, None], cr["beta_out"][:, None], cr["p"],
                                 cr["wl"], cr["wr"]], axis=1) if len(cr) else np.zeros((0, 36))
            arrays[f"cr{d}_int"] = ii.astype(np.int32)
            arrays[f"cr{d}_dbl"] = dd.astype(np.float64)
        arrays["u_final"] = ref["u_final"]
        nk = min(steps_kept(case), int(ref["summary"]["steps"]))
        arrays["u_steps"] = ref["u_steps"][:nk]
        sm = ref["summary"]
        arrays["summary"] = np.array([float(sm["steps"]), float(sm["t_final"]), float(sm["lmax"]),
                                      float(sm["l2"])] + [float(v) for v in sm["argmax"]])
        if "jumps" in ref:
            arrays["jumps"] = ref["jumps"]
        path = os.path.join(HERE, case["name"] + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"{case['name']}: {os.path.getsize(path) / 1024:.0f} KiB, steps {sm['steps']}, "
              f"lmax {sm['lmax']}")
    tmp.cleanup()


if __name__ == "__main__":
    main()
