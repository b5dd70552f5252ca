#!/usr/bin/env python3
"""adjustMiniStacks.py: datum adjustment of sequential mini-stack phases; same
flags as the reference script (python/adjustMiniStacks.py:17-52)."""
import argparse
import os


def cmdLineParser():
    parser = argparse.ArgumentParser(
        description='Adjusts mini-stack phase series with the datum adjustment phase.',
        formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument('-s', '--slcDir', type=str, dest='slcDir', required=True,
                        help='Input folder which contains vrt files of all slcs')
    parser.add_argument('-m', '--miniStackDir', type=str, dest='miniStackDir', required=True,
                        help='Input mini stack directory')
    parser.add_argument('-d', '--datumDir', type=str, dest='datumDir', required=True,
                        help='Input datum connection directory')
    parser.add_argument('-M', '--miniStackSize', type=int, dest='miniStackSize', required=True,
                        help='size of each miniStack')
    parser.add_argument('-o', '--outDir', type=str, dest='outDir', required=True, help='output directory')
    parser.add_argument('-t', '--timeSeries_vrt_file', type=str, dest='timeSeriesName', required=False,
                        help='output vrt file for time-series')
    parser.add_argument('-x', '--reference_x', type=int, dest='ref_x', required=False,
                        help='x coordinate of reference pixel')
    parser.add_argument('-y', '--reference_y', type=int, dest='ref_y', required=False,
                        help='y coordinate of reference pixel')
    parser.add_argument('--unwrapped', action='store_true', default=False,
                        help='adjusting unwrapped dataset. Default: False which means adjustment is done on wrapped data')
    return parser


def main(argv=None):
    inps = cmdLineParser().parse_args(argv)
    from fringe_amd import adjust
    os.makedirs(inps.outDir, exist_ok=True)
    dateList = adjust.get_dates(inps.slcDir)
    if inps.unwrapped:
        adjust.adjust_unwrapped(dateList, inps.miniStackDir, inps.datumDir, inps.outDir, inps.miniStackSize,
                                inps.timeSeriesName, inps.ref_x, inps.ref_y)
    else:
        adjust.adjust_wrapped(dateList, inps.miniStackDir, inps.datumDir, inps.outDir, inps.miniStackSize)
    return 0


if __name__ == '__main__':
    raise SystemExit(main())
