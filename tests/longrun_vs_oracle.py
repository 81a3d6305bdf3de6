"""The long-horizon FP32-mixed / FP64 comparison (sphera_amd.longrun) with the
FP64 oracle stepped beside the device: at every comparison the device FP64
state must equal the oracle's bit for bit, and FP32-mixed is measured
against the oracle (max and RMS per field).  A script, not a collected test
(2,000 oracle steps of the 576,000-particle case take about 5 minutes on 16
host threads):

    python tests/longrun_vs_oracle.py [--steps 2000] [--every 100] > longrun.md
"""
import argparse
import json
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))

from oracle.pyoracle import Oracle  # noqa: E402
from sphera_amd import longrun  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--dx", type=float, default=0.005)
    ap.add_argument("--rows-json", default=None, help="also write the rows as JSON to this file")
    a = ap.parse_args()
    tb, coord = longrun.make_case(a.dx)
    o = Oracle(tb, coord=coord)
    o.preloop()
    state = {"prev": 0}

    def extra(it, fp32, fp64):
        o.step(it - state["prev"])
        state["prev"] = it
        s = o.state()
        same = all(np.array_equal(fp64[f], s[f]) for f in longrun.FIELDS) and fp64["dt"] == s["dt"]
        e = longrun.field_errors(fp32, s)
        print(f"[longrun] step {it}: device FP64 bit-identical to the oracle: {same}", file=sys.stderr,
              flush=True)
        return {"fp64_equals_oracle": bool(same),
                **{"oracle_" + k: v for k, v in e.items() if not k.endswith("_argmax")}}

    rows = longrun.run(a.steps, a.every, a.dx, log=lambda s: print(s, file=sys.stderr, flush=True),
                       extra=extra)
    print(longrun.markdown(rows, a.dx))
    print()
    print("| step | device FP64 = oracle | " +
          " | ".join(f"FP32 vs oracle {f} max / rms" for f in longrun.FIELDS) + " |")
    print("|" + "---|" * (2 + len(longrun.FIELDS)))
    for r in rows:
        print(f"| {r['it']} | {'yes' if r['fp64_equals_oracle'] else 'NO'} | " +
              " | ".join(f"{r['oracle_' + f + '_max_rel']:.1e} / {r['oracle_' + f + '_rms_rel']:.1e}"
                         for f in longrun.FIELDS) + " |")
    if a.rows_json:
        with open(Path(a.rows_json), "w") as fh:
            json.dump(rows, fh)


if __name__ == "__main__":
    main()
