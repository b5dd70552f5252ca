#!/usr/bin/env python3
"""despeck.py: SHP-masked despeckling of SLC amplitude or a single-look
interferogram; same flags as the reference driver (src/despeck/despeck.py:6-32)."""
import argparse
import os


def cmdLineParser():
    parser = argparse.ArgumentParser(description='Despeckle SLC amplitude / single look interferogram',
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument('-i', '--input', type=str, dest='inputDS', required=True, help='Input GDAL SLC stack VRT')
    parser.add_argument('-o', '--output', type=str, dest='outputDS', required=True,
                        help='Output despeckled dataset')
    parser.add_argument('-w', '--wts', type=str, dest='wtsDS', default='',
                        help='Optional mask layer to speed up computation')
    parser.add_argument('-l', '--linesperblock', type=int, dest='linesPerBlock', default=64,
                        help='Quantum for block of lines')
    parser.add_argument('-r', '--ram', type=int, dest='memorySize', default=512, help='Memory in Mb to use')
    parser.add_argument('-x', '--xhalf', type=int, dest='halfWindowX', default=5, help='Half window size (range)')
    parser.add_argument('-y', '--yhalf', type=int, dest='halfWindowY', default=5, help='Half window size (azimuth)')
    parser.add_argument('-b', '--band', type=int, dest='bands', nargs='*', default=[],
                        help='Single band for SLC despeckling and 2 bands for interferogram')
    parser.add_argument('-c', '--corr', action='store_true', dest='cohFlag', default=False,
                        help='Replace amplitude for despecked interferogram with coherence')
    return parser


def runDespeck(inps):
    from fringe_amd.despecklib import Despeck
    aa = Despeck()
    aa.inputDS = inps.inputDS
    aa.weightsDS = inps.wtsDS
    aa.outputDS = inps.outputDS
    aa.blocksize = inps.linesPerBlock
    aa.memsize = inps.memorySize
    aa.halfWindowX = inps.halfWindowX
    aa.halfWindowY = inps.halfWindowY
    if len(inps.bands) == 1:
        aa.band1 = inps.bands[0]
        if inps.cohFlag:
            raise Exception('User requested coherence when requesting despeckling of SLC magnitude')
    elif len(inps.bands) == 2:
        aa.band1 = inps.bands[0]
        aa.band2 = inps.bands[1]
        aa.coherenceFlag = inps.cohFlag
    elif len(inps.bands) != 0:
        raise Exception('Despeck can handle one or two bands. More than two bands provided')
    return aa.run()


def main(argv=None):
    inps = cmdLineParser().parse_args(argv)
    outDir = os.path.dirname(os.path.abspath(inps.outputDS))
    os.makedirs(outDir, exist_ok=True)
    return runDespeck(inps)


if __name__ == '__main__':
    raise SystemExit(main())
