"""Command line: ``python -m maze_poisson_amd md INPUT.yml [-v] [--lib PATH]`` runs an
MD simulation from the reference's input schema (the counterpart of ``maze run md``,
``maze_poisson/cli/run.py:13-36``); ``python -m maze_poisson_amd convert OLD.csv DIR``
converts an old-schema input file."""
from __future__ import annotations

import argparse
import sys


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if argv and argv[0] == "convert":
        from .convert import main as convert_main
        return convert_main(argv[1:])
    ap = argparse.ArgumentParser(prog="python -m maze_poisson_amd")
    sub = ap.add_subparsers(dest="cmd", required=True)
    md = sub.add_parser("md", help="run MD from a YAML / JSON input file")
    md.add_argument("filename", metavar="YAML_INPUT_FILE")
    md.add_argument("-v", "--verbose", action="store_true")
    md.add_argument("--lib", default=None,
                    help="library exporting the MaZe ABI (default: the HIP library)")
    sub.add_parser("convert", help="convert an old-schema input CSV (see convert -h)")
    a = ap.parse_args(argv)
    from .solver import run_file
    lib = None
    if a.lib:
        from .capi import MazeLib
        lib = MazeLib(a.lib)
    try:
        s = run_file(a.filename, lib=lib, verbose=a.verbose)
    except ValueError as e:
        print(e, file=sys.stderr)
        return 2
    print(f"done: {s.mdv.N_steps} steps, {s.timers.get('steps_per_s', 0):.1f} steps/s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
