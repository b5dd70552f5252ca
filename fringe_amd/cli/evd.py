#!/usr/bin/env python3
"""evd.py: phase linking with the evd module (EVD / MLE / STBAS); same flags as
the reference driver (src/evd/evd.py:11-34), run on the GPU.  Extra flag:
--precision {mixed,fp64,mixed_bf16}."""
import argparse
import glob
import os

import numpy as np


def cmdLineParser(description='Perform MLE-based phase-linking on a stack of coregistered SLCs'):
    parser = argparse.ArgumentParser(description=description,
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument('-i', '--input', type=str, dest='inputDS', required=True, help='Input GDAL SLC stack VRT')
    parser.add_argument('-w', '--wts', type=str, dest='wtsDS', required=True, help='Input weights dataset')
    parser.add_argument('-o', '--output', type=str, dest='outputFolder', required=True,
                        help='Output folder with phase-linked SLCs')
    parser.add_argument('-l', '--linesperblock', type=int, dest='linesPerBlock', default=64,
                        help='Quantum for block of lines')
    parser.add_argument('-r', '--ram', type=int, dest='memorySize', default=2048, help='Memory in Mb to use')
    parser.add_argument('-x', '--xhalf', type=int, dest='halfWindowX', default=5, help='Half window size (range)')
    parser.add_argument('-y', '--yhalf', type=int, dest='halfWindowY', default=5, help='Half window size (azimuth)')
    parser.add_argument('-n', '--minneigh', type=int, dest='minNeighbors', default=5,
                        help='Minimum number of neighbors for computation')
    parser.add_argument('-m', '--method', type=str, dest='method', default='MLE',
                        help='Decomposition method to use - MLE / EVD / STBAS')
    parser.add_argument('-b', '--bandwidth', type=int, dest='bandWidth', default=-1,
                        help='Diagonal bandwidth for STBAS')
    parser.add_argument('--precision', type=str, default='mixed', choices=['mixed', 'fp64', 'mixed_bf16'],
                        help='GPU arithmetic: fp16-split MFMA Gram + FP32 eigensolver, or FP64 throughout')
    return parser


def run(inps, lib_cls):
    aa = lib_cls()
    aa.inputDS = inps.inputDS
    aa.weightsDS = inps.wtsDS
    aa.outputFolder = inps.outputFolder
    aa.outputCompressedSlcFolder = aa.outputFolder
    aa.compSlc = "compslc.bin"
    aa.blocksize = inps.linesPerBlock
    aa.memsize = inps.memorySize
    aa.halfWindowX = inps.halfWindowX
    aa.halfWindowY = inps.halfWindowY
    aa.minimumNeighbors = inps.minNeighbors
    aa.method = inps.method
    aa.bandWidth = inps.bandWidth
    aa.precision = inps.precision
    return aa.run()


def create_vrts(slc_dir):
    """A .vrt next to every output .slc (evd.py:69-76 does this with gdal.Translate)."""
    from fringe_amd import raster
    for f in glob.glob(os.path.join(slc_dir, "*.slc")):
        cols, rows, bands = raster.info(f)
        with open(f + ".vrt", "w") as fid:
            fid.write(raster.envi_vrt(f, cols, rows, bands, np.complex64))


def main(argv=None):
    from fringe_amd.evdlib import Evd
    inps = cmdLineParser().parse_args(argv)
    rc = run(inps, Evd)
    if rc == 0:
        create_vrts(inps.outputFolder)
    return rc


if __name__ == '__main__':
    raise SystemExit(main())
