"""Stack VRT assembly for mini-stacks, without GDAL.

Same folder layout and VRT content as src/sequential/Stack.py:36-188: every
SLC gets a raw VRT in <out>/slcs/<date>.vrt and the stack is
<out>/stack/stack.vrt with one SimpleSource band per SLC (SrcRect = bbox
for real SLCs, whole image for compressed SLCs) and Date metadata in domain
"slc".
"""
import glob
import os

import numpy as np

from . import raster


def _size(slc):
    vrt = slc + ".vrt"
    cols, rows, _ = raster.info(vrt if os.path.exists(vrt) else slc)
    return cols, rows


class Stack(object):
    def __init__(self, slcDir=None):
        self.slcDir = slcDir
        self.bbox = None

    def configure(self, outDir):
        os.makedirs(outDir, exist_ok=True)
        self.outSlcVrtDir = os.path.join(outDir, "slcs")
        self.outStackVrtDir = os.path.join(outDir, "stack")
        os.makedirs(self.outSlcVrtDir, exist_ok=True)
        os.makedirs(self.outStackVrtDir, exist_ok=True)

    def gatherSLCs(self):
        self.slcList = glob.glob(os.path.join(self.slcDir, '*/*.slc'))
        if len(self.slcList) == 0:
            self.slcList = glob.glob(os.path.join(self.slcDir, '*/*.slc.full'))
        self.slcList.sort()
        self.size = len(self.slcList)
        self.applyBbox = [True] * self.size

    def getSize(self):
        self.size = len(self.slcList)

    def getDates(self):
        self.dateList = [os.path.basename(os.path.dirname(slc)) for slc in self.slcList]

    def get_x_y_offsets(self, ind):
        width, height = _size(self.slcList[ind])
        ymin, ymax, xmin, xmax = 0, height, 0, width
        if self.bbox and self.applyBbox[ind]:
            ymin, ymax, xmin, xmax = self.bbox
        return width, height, xmin, ymin, xmax - xmin, ymax - ymin

    def writeStackVRT(self):
        dates = []
        for slc in self.slcList:
            width, height = _size(slc)
            date = os.path.basename(os.path.dirname(slc))
            with open(os.path.join(self.outSlcVrtDir, f'{date}.vrt'), 'w') as fid:
                fid.write(raster.raw_vrt(os.path.abspath(slc), width, height, np.complex64))
            dates.append(date)
        self.stackVRT = os.path.join(self.outStackVrtDir, 'stack.vrt')
        _, _, _, _, xsize, ysize = self.get_x_y_offsets(-1)
        out = [f'<VRTDataset rasterXSize="{xsize}" rasterYSize="{ysize}">']
        for ind, date in enumerate(dates):
            width, height, xmin, ymin, xs, ys = self.get_x_y_offsets(ind)
            path = os.path.abspath(os.path.join(self.outSlcVrtDir, date + '.vrt'))
            out.append(f'''    <VRTRasterBand dataType="CFloat32" band="{ind + 1}">
        <SimpleSource>
            <SourceFilename>{path}</SourceFilename>
            <SourceBand>1</SourceBand>
            <SourceProperties RasterXSize="{width}" RasterYSize="{height}" DataType="CFloat32"/>
            <SrcRect xOff="{xmin}" yOff="{ymin}" xSize="{xs}" ySize="{ys}"/>
            <DstRect xOff="0" yOff="0" xSize="{xs}" ySize="{ys}"/>
        </SimpleSource>
        <Metadata domain="slc">
            <MDI key="Date">{date}</MDI>
            <MDI key="Wavelength">0.03</MDI>
            <MDI key="AcquisitionTime">{date}</MDI>
        </Metadata>
    </VRTRasterBand>''')
        out.append('</VRTDataset>')
        with open(self.stackVRT, 'w') as fid:
            fid.write('\n'.join(out))


class MiniStack(Stack):
    def updateMiniStack(self, compressedSlcDir):
        # Stack.py:183-188 prepends every compressed SLC found so far.  The
        # reference takes them in glob order; they are sorted here so the
        # band order is reproducible (the estimator's reference index, the
        # first real SLC, does not depend on it).
        compSlcList = sorted(glob.glob(os.path.join(compressedSlcDir, '*/*.slc')))
        self.slcList = compSlcList + self.slcList
        self.applyBbox = [False] * len(compSlcList) + self.applyBbox
