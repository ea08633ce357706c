// idwOperator.cpp -- see idwOperator.h.
#include "idwOperator.h"

#include <climits>
#include <cmath>
#include <vector>

#include "../../../include/pargo.h"
#include "../../gpro/shapefile.h"

bool idwMaskForDevice(RasterLayer<int> *mask, std::vector<int> &buf, int &rows, int &cols)
{
    const CellSpace<int> *mc = mask ? mask->cellSpace() : nullptr;
    if (!mc || mc->empty() || !mask->metaData()) {
        cerr << "[ERROR] IDW: the mask layer holds no data" << endl;
        return false;
    }
    const MetaData &m = *mask->metaData();
    const int local = m._localworkBR.maxIRow() + 1;  // rows the lookup can reach
    rows = m._glbDims.nRows();
    cols = mc->nCols();
    if (local < 1 || local > mc->nRows() || rows < local || m._MBR.minIRow() != 0) {
        cerr << "[ERROR] IDW: unsupported mask band (rows " << m._MBR.minIRow() << ".." << m._MBR.maxIRow() << " of "
             << rows << ")" << endl;
        return false;
    }
    buf.resize((size_t)rows * cols);
    for (int r = 0; r < rows; r++) {
        const int *src = (*mc)[std::min(r, local - 1)];
        std::copy(src, src + cols, buf.begin() + (size_t)r * cols);
    }
    return true;
}

IDWOperator::~IDWOperator()
{
    if (_plan) pargo_idw_plan_free(Application::device(), _plan);
}

int IDWOperator::readSampleNums(const char *filename, char **pSpatialRefWkt)
{
    shapefile::Points pts;
    std::string err;
    if (!shapefile::read_points(filename, -1, pts, err)) {
        cerr << err << endl;
        return 0;
    }
    _wkt = pts.projection_wkt;
    if (pSpatialRefWkt) *pSpatialRefWkt = const_cast<char *>(_wkt.c_str());
    _sample_nums = (int)pts.x.size();
    return _sample_nums;
}

bool IDWOperator::readSamples(const char *filename, int fieldIdx, char **pSpatialRefWkt,
                              vector<SamplePoint> &samples)
{
    shapefile::Points pts;
    std::string err;
    if (!shapefile::read_points(filename, fieldIdx, pts, err)) {
        cerr << err << endl;
        return false;
    }
    if (pts.x.empty()) {
        cerr << "[ERROR] no samples in " << filename << endl;
        return false;
    }
    _wkt = pts.projection_wkt;
    if (pSpatialRefWkt) *pSpatialRefWkt = const_cast<char *>(_wkt.c_str());
    for (size_t i = 0; i < pts.x.size(); i++) {
        SamplePoint s;
        s.x = pts.x[i], s.y = pts.y[i], s.value = pts.value[i];
        samples.push_back(s);
    }
    double ext[4];
    if (!deviceOk(pargo_idw_extent(pts.x.data(), pts.y.data(), (int)pts.x.size(), _cellSize, ext, &_nRows, &_nCols),
                  "pargo_idw_extent"))
        return false;
    _glb_extent.minX = ext[0], _glb_extent.maxX = ext[1], _glb_extent.minY = ext[2], _glb_extent.maxY = ext[3];
    return true;
}

void IDWOperator::creatSampleBlocks(vector<SamplePoint> &samples)
{
    // host buckets in file order (idwOperator.cpp:152-169), for getSampleBlocks()
    _blockRows = (int)std::ceil((_glb_extent.maxY - _glb_extent.minY) / _blockSize);
    _blockCols = (int)std::ceil((_glb_extent.maxX - _glb_extent.minX) / _blockSize);
    _pSampleBlocks.assign((size_t)std::max(0, _blockRows) * std::max(0, _blockCols), SampleBlock());
    std::vector<double> xs, ys, vs;
    for (const SamplePoint &s : samples) {
        const int iCol = getBlockColIndexByCoord(s.x), iRow = getBlockRowIndexByCoord(s.y);
        const long long index = (long long)iRow * _blockCols + iCol;
        if (iCol >= 0 && iRow >= 0 && iCol < _blockCols && index < (long long)_pSampleBlocks.size())
            _pSampleBlocks[index].samplePoints.push_back(s);
        xs.push_back(s.x), ys.push_back(s.y), vs.push_back(s.value);
    }
    pargo_ctx *ctx = Application::device();
    if (!ctx) return;
    if (_plan) pargo_idw_plan_free(ctx, _plan), _plan = nullptr;
    deviceOk(pargo_idw_prepare(ctx, xs.data(), ys.data(), vs.data(), (int)xs.size(), _cellSize, _blockSize, &_plan),
             "pargo_idw_prepare");
}

bool IDWOperator::setLayerMeta(RasterLayer<double> &layerD, char **pSpatialRefWkt)
{
    MetaData m;
    m.noData = _noData;
    m.row = _nRows;
    m.column = _nCols;
    m._glbDims = SpaceDims(_nRows, _nCols);
    m.cellSize = _cellSize;
    m.format = "GTiff";
    m._domDcmpType = ROWWISE_DCMP;
    m.dataType = DT_Float64;
    m.projection = (pSpatialRefWkt && *pSpatialRefWkt) ? *pSpatialRefWkt : _wkt;
    m.pTransform[0] = _glb_extent.minX;
    m.pTransform[1] = _cellSize;
    m.pTransform[2] = 0;
    m.pTransform[3] = _glb_extent.maxY;
    m.pTransform[4] = 0;
    m.pTransform[5] = -_cellSize;
    m._localdims = m._glbDims;
    m._MBR = CoordBR(CellCoord(0, 0), CellCoord(_nRows - 1, _nCols - 1));
    m._localworkBR = m._MBR;
    layerD.setMetaData(m);
    return true;
}

void IDWOperator::initIdwLayerGlobalInfo(RasterLayer<double> &layerD, char **pSpatialRefWkt)
{
    setLayerMeta(layerD, pSpatialRefWkt);
}

void IDWOperator::idwLayer(RasterLayer<double> &layerD, char **pSpatialRefWkt, DomDcmpType dcmpType)
{
    setLayerMeta(layerD, pSpatialRefWkt);
    if (!layerD.nbrhood()) layerD.newLocalNbrhood();
    layerD.layerDcmp(dcmpType);
    _pIDWLayer = &layerD;
    Configure(_pIDWLayer, false);
}

void IDWOperator::idwLayer(RasterLayer<double> &layerD, char **pSpatialRefWkt, CoordBR &subWorkBR)
{
    idwLayer(layerD, pSpatialRefWkt, ROWWISE_DCMP);
    layerD._pMetaData->_localworkBR = subWorkBR;
}

bool IDWOperator::Operator(const CellCoord &, bool)
{
    cerr << "[ERROR] IDWOperator runs on the GPU only (WorkDevice)" << endl;
    return false;
}

bool IDWOperator::WorkDevice(const CoordBR &workBR, bool &ok)
{
    ok = false;
    if (!_pIDWLayer || !_pIDWLayer->cellSpace()) {
        cerr << "[ERROR] IDWOperator: idwLayer() must be set before Run()" << endl;
        return true;
    }
    if (!_plan) {
        cerr << "[ERROR] IDWOperator: creatSampleBlocks() must succeed before Run()" << endl;
        return true;
    }
    const int rows = _nRows, cols = _nCols;
    if (!workBR.valid(SpaceDims(rows, cols))) {
        cerr << "[ERROR] IDWOperator: work rectangle (" << workBR << ") outside the " << rows << " x " << cols
             << " grid" << endl;
        return true;
    }
    pargo_ctx *ctx = Application::device();
    if (!ctx) return true;
    const size_t n = (size_t)rows * cols;
    void *d_out = nullptr, *d_mask = nullptr, *d_cand = nullptr;
    int mrows = 0, mcols = 0, mnodata = 0;
    bool good = deviceOk(pargo_alloc(ctx, n * sizeof(float), &d_out), "pargo_alloc(idw)") &&
                deviceOk(pargo_alloc(ctx, sizeof(unsigned long long), &d_cand), "pargo_alloc(count)");
    std::vector<int> mbuf;
    if (good && _pMaskLayer) {
        // int maskNoData = metaData()->noData (idwOperator.cpp:450)
        const double nd = _pMaskLayer->metaData() ? _pMaskLayer->metaData()->noData : 0;
        mnodata = (nd >= INT_MIN && nd <= INT_MAX) ? (int)nd : 0;
        good = idwMaskForDevice(_pMaskLayer, mbuf, mrows, mcols) &&
               deviceOk(pargo_alloc(ctx, mbuf.size() * sizeof(int), &d_mask), "pargo_alloc(mask)") &&
               deviceOk(pargo_h2d(ctx, d_mask, mbuf.data(), mbuf.size() * sizeof(int)), "pargo_h2d(mask)");
    }
    std::vector<float> host(good ? n : 0);
    std::vector<double> work;
    void *d_work = nullptr;
    if (_writePreExpLoad) {
        // The reference's preliminary experiment replaces each value by the
        // cell's CPU time (idwOperator.cpp:455-458,482-484); the device
        // records the samples each cell's search examined instead.
        work.resize(good ? n : 0);
        good = good && deviceOk(pargo_alloc(ctx, n * sizeof(double), &d_work), "pargo_alloc(work)") &&
               deviceOk(pargo_idw_work(ctx, _plan, _nbrPoints, _idw_power, _idw_buffer, (const int32_t *)d_mask,
                                       mrows, mcols, mnodata, (float *)d_out, (double *)d_work),
                        "pargo_idw_work") &&
               deviceOk(pargo_d2h(ctx, work.data(), d_work, n * sizeof(double)), "pargo_d2h(work)");
    } else {
        good = good && deviceOk(pargo_idw_run(ctx, _plan, _nbrPoints, _idw_power, _idw_buffer,
                                              (const int32_t *)d_mask, mrows, mcols, mnodata, (float *)d_out,
                                              (unsigned long long *)d_cand),
                                "pargo_idw_run") &&
               deviceOk(pargo_d2h(ctx, &_candidates, d_cand, sizeof(_candidates)), "pargo_d2h(count)");
    }
    good = good && deviceOk(pargo_d2h(ctx, host.data(), d_out, n * sizeof(float)), "pargo_d2h(idw)") &&
           deviceOk(pargo_sync(ctx), "pargo_sync");
    for (void *p : {d_out, d_mask, d_cand, d_work})
        if (p) pargo_free(ctx, p);
    if (!good) return true;
    CellSpace<double> &idwL = *_pIDWLayer->cellSpace();
    for (int r = workBR.minIRow(); r <= workBR.maxIRow(); r++)
        for (int c = workBR.minICol(); c <= workBR.maxICol(); c++)
            idwL[r][c] = _writePreExpLoad ? work[(size_t)r * cols + c] : (double)host[(size_t)r * cols + c];
    flag = true;
    ok = true;
    return true;
}
