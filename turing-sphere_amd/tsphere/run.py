"""Command-line run of one model (the reference's `npm run bench --`,
scripts/bench.ts): ``python -m tsphere.run --preset schnak-spots --lmax 63
--steps 2000 --seed 1 [--json] [--dump-state FILE] [--batch 16]`` plus every
model parameter as ``--<key> <value>``.  Prints the run's JSON record with
--json, else a short report; the exit status is 1 if the state went
non-finite."""
from __future__ import annotations

import json
import sys

from .session import StateDigest, _num, format_command, format_digest, parse_args, run


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    want_json = "--json" in argv
    argv = [a for a in argv if a != "--json"]
    dump, batch, legendre = None, 16, "fp32"
    rest = []
    i = 0
    while i < len(argv):
        a = argv[i]
        if a in ("--dump-state", "--batch", "--legendre") and i + 1 < len(argv):
            v = argv[i + 1]
            if a == "--dump-state":
                dump = v
            elif a == "--batch":
                batch = int(v)
            else:
                legendre = v
            i += 2
            continue
        rest.append(a)
        i += 1
    try:
        spec = parse_args(rest)
    except ValueError as e:
        print(f"error: {e}", file=sys.stderr)
        return 2
    rec = run(spec, batch=batch, dump_state=dump, legendre=legendre)
    if want_json:
        print(json.dumps(rec, indent=2))
    else:
        thr = rec["throughput"]
        print(f"  {format_command(spec)}")
        print(f"  grid      lmax {rec['grid']['lmax']} . {rec['grid']['nlat']}x"
              f"{rec['grid']['nphi']} . nlm {rec['grid']['nlm']}")
        print(f"  compiled  {rec['compiled']['opsPerStep']} GPU ops/step "
              f"({rec['compiled']['kernels']} generated kernels)")
        print(f"  {thr['msPerStep']:.3f} ms/step   {thr['stepsPerSec']:.1f} steps/s "
              f"(batches of {thr['batch']})")
        if rec["digest"]:
            print(f"  state after {spec.steps} steps from seed {_num(spec.seed)}:")
            print(f"    {format_digest(StateDigest(**rec['digest']))}")
    return 0 if rec["state"]["finite"] else 1


if __name__ == "__main__":
    sys.exit(main())
