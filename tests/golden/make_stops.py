"""Record where the reference build stops on its own (exit(0) in the set-up of
its tangential-jump approximations or in its first step) for a set of
geometries, into tests/golden/ref_stops.json.

Run in the build container, where `make -C oracle ref` has built
oracle/_ref/mts_ref from the reference's sources:

    python tests/golden/make_stops.py

Each case runs one LOD-CN step with MIB-L1 and keeps the reference's exit
status, whether it completed the step and the last line it printed (its
exit message when it stopped).  Data only: no reference source is stored.
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "mts_ref")
R0 = 0.5 * (1 - 2 ** -20)
E1 = 1 - 2 ** -20

CASES = [
    dict(name="sphere32", n=32, surface=("ellipsoid", 1, 1, 1, R0)),
    dict(name="sphere64", n=64, surface=("ellipsoid", 1, 1, 1, R0)),
    dict(name="sphere128", n=128, surface=("ellipsoid", 1, 1, 1, R0)),
    dict(name="sphere256", n=256, surface=("ellipsoid", 1, 1, 1, R0)),
    dict(name="ellipsoid32", n=32, surface=("ellipsoid", 0.7, 0.5, 0.4, E1)),
    dict(name="ellipsoid64", n=64, surface=("ellipsoid", 0.7, 0.5, 0.4, E1)),
    dict(name="ellipsoid128", n=128, surface=("ellipsoid", 0.7, 0.5, 0.4, E1)),
    dict(name="ellipsoid256", n=256, surface=("ellipsoid", 0.7, 0.5, 0.4, E1)),
    dict(name="torus22", n=22, surface=("torus", 0.5, 0.17)),
    dict(name="torus48", n=48, surface=("torus", 0.5, 0.2)),
    dict(name="torus96", n=96, surface=("torus", 0.5, 0.2)),
    dict(name="bumpy26", n=26, surface=("bumpy", R0, 0.1, 5, 3)),
    dict(name="bumpy64", n=64, surface=("bumpy", R0, 0.1, 5, 3)),
]


def run(case):
    kind, *p = case["surface"]
    surf = f"{kind}:" + ",".join(repr(float(v)) for v in p)
    with tempfile.TemporaryDirectory() as out:
        cmd = [REF, "--method", "lod-cn", "--mib", "1", "--n", str(case["n"]), "--lo", "-1",
               "--hi", "1", "--dt", repr(2.0 / case["n"]), "--T", repr(2.0 / case["n"]),
               "--surface", surf, "--beta", "const:10,1", "--out", out, "--max-steps", "1",
               "--no-dump"]
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        summary = os.path.exists(os.path.join(out, "summary.txt"))
    lines = [l.strip() for l in p.stdout.splitlines() if l.strip()]
    return dict(case, returncode=p.returncode, completed=summary,
                last_line=lines[-1] if lines else "")


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/mts_ref is not built (make -C oracle ref)")
    res = []
    for c in CASES:
        r = run(c)
        print(r["name"], r["returncode"], r["completed"], r["last_line"], flush=True)
        res.append(r)
    with open(os.path.join(HERE, "ref_stops.json"), "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
