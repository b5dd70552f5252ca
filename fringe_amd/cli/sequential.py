#!/usr/bin/env python3
"""sequential.py: mini-stack sequential phase linking; same flags as the
reference driver (src/sequential/sequential.py:10-42)."""
import argparse
import os


def cmdLineParser():
    parser = argparse.ArgumentParser(description='Perform MLE-based phase-linking on a stack of coregistered SLCs',
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument('-i', '--inDir', type=str, dest='inputDir', required=True,
                        help='Input folder which contains folders for each SLC')
    parser.add_argument('-w', '--weight_dataset', type=str, dest='weightDS', required=True, help='Input weights dataset')
    parser.add_argument('-o', '--outDir', type=str, dest='outputDir', required=True, help='Output folder')
    parser.add_argument('-l', '--linesperblock', type=int, dest='linesPerBlock', default=64,
                        help='Quantum for block of lines')
    parser.add_argument('-r', '--ram', type=int, dest='memorySize', default=2048, help='Memory in Mb to use')
    parser.add_argument('-x', '--xhalf', type=int, dest='halfWindowX', default=29, help='Half window size (range)')
    parser.add_argument('-y', '--yhalf', type=int, dest='halfWindowY', default=9, help='Half window size (azimuth)')
    parser.add_argument('-m', '--minneigh', type=int, dest='minNeighbors', default=5,
                        help='Minimum number of neighbors for computation')
    parser.add_argument('-b', '--bbox', dest='bbox', nargs='+', type=str, default=None,
                        help='bounding box : minLine maxLine minPixel maxPixel \nORcoreg_stack/slcs_base.vrt')
    parser.add_argument('-s', '--mini_stack_size', type=int, dest='miniStackSize', default=10, help='mini stack size')
    parser.add_argument('-f', '--force', dest='forceprocessing', action='store_true', default=False,
                        help='Force reprocessing')
    parser.add_argument('--precision', type=str, default='mixed', choices=['mixed', 'fp64', 'mixed_bf16'])
    return parser


def vrt_file2bbox(vrt_file):
    """bbox (ymin, ymax, xmin, xmax) from the SrcRect of a stack VRT (sequential.py:45-62)."""
    import xml.etree.ElementTree as ET
    root = ET.parse(vrt_file).getroot()
    tag = root.find('VRTRasterBand/SimpleSource/SrcRect')
    xmin, ymin = int(tag.get('xOff')), int(tag.get('yOff'))
    return (ymin, ymin + int(tag.get('ySize')), xmin, xmin + int(tag.get('xSize')))


def main(argv=None):
    from fringe_amd.sequential import run_sequential
    inps = cmdLineParser().parse_args(argv)
    bbox = None
    if inps.bbox is not None:
        bbox = vrt_file2bbox(inps.bbox[0]) if os.path.isfile(inps.bbox[0]) else tuple(int(i) for i in inps.bbox)
    res = run_sequential(inps.inputDir, inps.weightDS, inps.outputDir, inps.miniStackSize, bbox,
                         inps.linesPerBlock, inps.memorySize, inps.halfWindowX, inps.halfWindowY,
                         inps.minNeighbors, inps.forceprocessing, inps.precision)
    return max((rc for _, rc in res), default=0)


if __name__ == '__main__':
    raise SystemExit(main())
