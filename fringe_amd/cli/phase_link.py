#!/usr/bin/env python3
"""phase_link.py: EMI phase linking with EVD fallback; same flags as the
reference driver (src/phase_link/phase_link.py:11-34), run on the GPU."""
from fringe_amd.cli.evd import cmdLineParser, create_vrts, run


def main(argv=None):
    from fringe_amd.phase_linklib import Phaselink
    inps = cmdLineParser().parse_args(argv)
    rc = run(inps, Phaselink)
    if rc == 0:
        create_vrts(inps.outputFolder)
    return rc


if __name__ == '__main__':
    raise SystemExit(main())
