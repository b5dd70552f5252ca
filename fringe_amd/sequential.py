"""Sequential mini-stack estimator (src/sequential/sequential.py:144-255).

The stack is split into mini-stacks of ``miniStackSize`` dates.  Mini-stack k
is [compressed SLCs of mini-stacks 1..k-1] + [its own SLCs], phase-linked by
the evd module (MLE by default, evd.hpp:66) with miniStackCount = k, so the
reference date is its first real SLC and only real SLCs are compressed
(evd.cpp:738-762).  Its compressed SLC is written to
compressedSlc/<lastDate>/<lastDate>.slc.  The datum connection then links the
stack of compressed SLCs with miniStackCount = 1.  A mini-stack whose output
folder exists is skipped unless force is set (sequential.py:206-207).
"""
import os

import numpy as np

from . import raster
from .evdlib import Evd
from .stack import MiniStack, Stack


def _run_evd(args, inputDataset, weightDS, outDir, miniStackCount, compressedSlcDir=None,
             compressedSlcName=None):
    aa = Evd()
    aa.inputDS = inputDataset
    aa.weightsDS = weightDS
    aa.outputFolder = outDir
    aa.miniStackCount = miniStackCount
    aa.blocksize = args["linesPerBlock"]
    aa.memsize = args["memorySize"]
    aa.halfWindowX = args["halfWindowX"]
    aa.halfWindowY = args["halfWindowY"]
    aa.minimumNeighbors = args["minNeighbors"]
    aa.method = args.get("method", "MLE")
    aa.precision = args.get("precision", "mixed")
    aa.verbose = args.get("verbose", True)
    aa.outputCompressedSlcFolder = compressedSlcDir if compressedSlcDir else aa.outputFolder
    aa.compSlc = compressedSlcName if compressedSlcName else "compslc.bin"
    return aa.run()


def run_sequential(inputDir, weightDS, outputDir, miniStackSize=10, bbox=None, linesPerBlock=64,
                   memorySize=2048, halfWindowX=29, halfWindowY=9, minNeighbors=5, force=False,
                   precision="mixed", method="MLE", verbose=True, mini_stack_sizes=None):
    """Run all mini-stacks and the datum connection; returns a list of
    (folder, return code).  mini_stack_sizes (optional) gives uneven
    mini-stack lengths instead of a fixed miniStackSize."""
    args = dict(linesPerBlock=linesPerBlock, memorySize=memorySize, halfWindowX=halfWindowX,
                halfWindowY=halfWindowY, minNeighbors=minNeighbors, precision=precision,
                method=method, verbose=verbose)
    outputDir = os.path.abspath(outputDir)
    compressedSlcDir = os.path.join(outputDir, "compressedSlc")
    os.makedirs(compressedSlcDir, exist_ok=True)
    stack = Stack(inputDir)
    stack.bbox = bbox
    stack.gatherSLCs()
    stack.getDates()
    stack.configure(os.path.join(outputDir, "fullStack"))
    stack.writeStackVRT()
    results = []
    miniStackCount = 0
    indStart = 0
    k = 0
    while indStart < stack.size:
        miniStackCount += 1
        size = miniStackSize if mini_stack_sizes is None else mini_stack_sizes[k % len(mini_stack_sizes)]
        k += 1
        indEnd = min(indStart + size, stack.size)
        startDate, endDate = stack.dateList[indStart], stack.dateList[indEnd - 1]
        outDir = os.path.join(outputDir, "miniStacks", startDate + "_" + endDate)
        if os.path.isdir(outDir) and not force:
            if verbose:
                print(f'{outDir} looks like it has already been processed. Skipping ... ')
        else:
            miniStack = MiniStack()
            miniStack.slcList = stack.slcList[indStart:indEnd]
            miniStack.getSize()
            miniStack.applyBbox = [True] * miniStack.size
            miniStack.updateMiniStack(compressedSlcDir)
            miniStack.bbox = bbox
            miniStack.getDates()
            miniStack.configure(outDir)
            miniStack.writeStackVRT()
            evdDir = os.path.join(outDir, "EVD")
            name = miniStack.dateList[-1] + ".slc"
            compSlcDir = os.path.join(compressedSlcDir, miniStack.dateList[-1])
            os.makedirs(compSlcDir, exist_ok=True)
            rc = _run_evd(args, miniStack.stackVRT, weightDS, evdDir, miniStackCount, compSlcDir, name)
            results.append((evdDir, rc))
            comp = os.path.join(compSlcDir, name)
            if rc == 0:
                cols, rows, bands = raster.info(comp)
                with open(comp + ".vrt", "w") as fid:   # sequential.py:231 (gdal_translate -of VRT)
                    fid.write(raster.envi_vrt(comp, cols, rows, bands, np.complex64))
        indStart = indEnd
    # datum connection (sequential.py:246-254)
    outDir = os.path.join(outputDir, "Datum_connection")
    comp = Stack(compressedSlcDir)
    comp.gatherSLCs()
    comp.bbox = None
    comp.applyBbox = [False] * comp.size
    comp.getDates()
    comp.configure(outDir)
    comp.writeStackVRT()
    rc = _run_evd(args, comp.stackVRT, weightDS, os.path.join(outDir, "EVD"), 1)
    results.append((os.path.join(outDir, "EVD"), rc))
    return results
