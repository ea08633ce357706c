// rasterLayer.h -- raster layer of the gpro interface (gpro/rasterLayer.h of
// the reference: RasterLayer<elemType> :41-236).
//
// A layer couples a host CellSpace, an optional Neighborhood and a MetaData,
// and (double layers) a device-resident copy of its cells (deviceRaster.h):
// device operators read and leave their data there, and the host CellSpace
// is brought up to date only when host code asks for it through
// cellSpace() or writes the file.  Non-const cellSpace() access counts as a
// host write and retires the device copy.
// Files are GeoTIFFs read and written by the GDAL-free codec in geotiff.h.
// One process owns the whole raster (this build runs one process per GPU),
// so every decomposition type reads all rows; the work rectangle is the set
// of cells whose neighborhood lies inside the raster, as in the reference.
#pragma once

#include <fstream>
#include <memory>
#include <string>
#include <type_traits>

#include "cellSpace.h"
#include "deviceRaster.h"
#include "geotiff.h"
#include "metaData.h"
#include "neighborhood.h"

namespace GPRO {

template <class elemType>
class RasterLayer {
  public:
    RasterLayer() : _strLayerName("Untitled") {}
    explicit RasterLayer(const string &layerName) : _strLayerName(layerName) {}
    RasterLayer(const RasterLayer &rhs) { *this = rhs; }
    RasterLayer &operator=(const RasterLayer &rhs)
    {
        if (this == &rhs) return *this;
        _strLayerName = rhs._strLayerName;
        _nLayerID = rhs._nLayerID;
        rhs.syncHost();
        _dev.reset();
        _pCellSpace.reset(rhs._pCellSpace ? new CellSpace<elemType>(*rhs._pCellSpace) : nullptr);
        _pNbrhood.reset(rhs._pNbrhood ? new Neighborhood<elemType>(*rhs._pNbrhood) : nullptr);
        _meta.reset(rhs._meta ? new MetaData(*rhs._meta) : nullptr);
        _pMetaData = _meta.get();
        return *this;
    }
    virtual ~RasterLayer() {}

    const string &name() const { return _strLayerName; }
    void name(const string &n) { _strLayerName = n; }
    int id() const { return _nLayerID; }
    void id(int i) { _nLayerID = i; }

    CellSpace<elemType> *cellSpace()
    {
        syncHost();
        retireDevice();  // the caller may write the cells
        return _pCellSpace.get();
    }
    const CellSpace<elemType> *cellSpace() const
    {
        syncHost();
        return _pCellSpace.get();
    }
    Neighborhood<elemType> *nbrhood() { return _pNbrhood.get(); }
    const Neighborhood<elemType> *nbrhood() const { return _pNbrhood.get(); }
    MetaData *metaData() { return _pMetaData; }
    const MetaData *metaData() const { return _pMetaData; }

    void cleanCellSpace() { _pCellSpace.reset(), _dev.reset(); }
    void cleanNbrhood() { _pNbrhood.reset(); }
    void cleanMetaData() { _meta.reset(), _pMetaData = nullptr; }
    bool hasCellSpace() const { return (bool)_pCellSpace; }
    bool hasNbrhood() const { return (bool)_pNbrhood; }
    bool hasMetaData() const { return _pMetaData != nullptr; }

    bool newCellSpace() { return retireDevice(), _pCellSpace.reset(new CellSpace<elemType>()), true; }
    bool newCellSpace(const SpaceDims &d) { return retireDevice(), _pCellSpace.reset(new CellSpace<elemType>(d)), true; }
    bool newCellSpace(const SpaceDims &d, const elemType &v)
    {
        return retireDevice(), _pCellSpace.reset(new CellSpace<elemType>(d, v)), true;
    }
    bool newCellSpace(int r, int c) { return newCellSpace(SpaceDims(r, c)); }
    bool newCellSpace(int r, int c, const elemType &v) { return newCellSpace(SpaceDims(r, c), v); }

    bool newNbrhood() { return _pNbrhood.reset(new Neighborhood<elemType>()), true; }
    bool newLocalNbrhood() { return _pNbrhood.reset(new Neighborhood<elemType>(true)), true; }
    bool newNbrhood(const vector<CellCoord> &coords, double weight = 1.0)
    {
        _pNbrhood.reset(new Neighborhood<elemType>());
        return _pNbrhood->fromVect(coords, weight);
    }
    bool newNbrhood(const vector<CellCoord> &coords, const vector<double> &weights)
    {
        _pNbrhood.reset(new Neighborhood<elemType>());
        return _pNbrhood->fromVects(coords, weights);
    }
    bool newRectangleNbrhood(int radius)
    {
        vector<CellCoord> coords;
        for (int r = -radius; r <= radius; r++)
            for (int c = -radius; c <= radius; c++) coords.push_back(CellCoord(r, c));
        return newNbrhood(coords, 1.0);
    }
    template <class T2>
    bool newNbrhood(const Neighborhood<T2> &nbr)
    {
        _pNbrhood.reset(new Neighborhood<elemType>(nbr));
        return true;
    }

    // Reads a .nbr file (count, then "iRow iCol weight" triples).
    bool readNeighborhood(const char *neighborfile)
    {
        newNbrhood();
        std::ifstream is(neighborfile);
        if (!is) {
            cerr << "[ERROR] cannot open neighborhood file " << neighborfile << endl;
            return false;
        }
        is >> *_pNbrhood;
        if (_pNbrhood->empty()) {
            cerr << "[ERROR] no neighborhood read from " << neighborfile << endl;
            return false;
        }
        return true;
    }
    bool readNeighborhoodSerial(const char *neighborfile) { return readNeighborhood(neighborfile); }

    // Reads band 1 of a GeoTIFF into this layer (rasterLayer.h:882-903).
    bool readFile(const char *inputfile, DomDcmpType dcmpType = NON_DCMP)
    {
        std::vector<double> values;
        if (!readFileInfo(inputfile, dcmpType, &values)) return false;
        if (!layerDcmp(dcmpType)) return false;
        retireDevice();
        elemType *dst = _pCellSpace->data();
        for (size_t i = 0; i < values.size(); i++) dst[i] = (elemType)values[i];
        return true;
    }
    // Reads the rows of subWorkBR plus the neighborhood's halo
    // (rasterLayer.h:927-955, layerDcmp :855-878): the cell space holds
    // rows [subWorkBR.minIRow() + nbr.minIRow(), min(subWorkBR.maxIRow(),
    // rows - 1) + nbr.maxIRow()] and the work rectangle is local to them.
    bool readFile(const char *inputfile, const CoordBR &subWorkBR, DomDcmpType dcmpType = NON_DCMP)
    {
        std::vector<double> values;
        if (!readFileInfo(inputfile, dcmpType, &values)) return false;
        if (!layerDcmp(subWorkBR)) return false;
        const int r0 = _pMetaData->_MBR.minIRow(), cols = _pMetaData->_glbDims.nCols();
        retireDevice();
        elemType *dst = _pCellSpace->data();
        const size_t n = (size_t)_pMetaData->_localdims.nRows() * cols;
        const double *src = values.data() + (size_t)r0 * cols;
        for (size_t i = 0; i < n; i++) dst[i] = (elemType)src[i];
        return true;
    }
    bool readGlobalFileSerial(const char *inputfile) { return readFile(inputfile, NON_DCMP); }

    // Sets the rectangles of a one-process decomposition and allocates the
    // cell space (rasterLayer.h:816-854).
    bool layerDcmp(DomDcmpType dcmpType)
    {
        if (!_pMetaData) return false;
        if (dcmpType == BLOCK_DCMP) {
            cerr << "RasterLayer::layerDcmp: block decomposition is not supported" << endl;
            return false;
        }
        MetaData &m = *_pMetaData;
        m._domDcmpType = dcmpType;
        m.myrank = 0;
        m.processor_number = 1;
        m._localdims = m._glbDims;
        m._MBR = CoordBR(CellCoord(0, 0), CellCoord(m._glbDims.nRows() - 1, m._glbDims.nCols() - 1));
        if (_pNbrhood && !_pNbrhood->empty()) {
            if (!_pNbrhood->calcWorkBR(m._localworkBR, m._glbDims)) return false;
        } else {
            m._localworkBR = m._MBR;
        }
        newCellSpace(m._localdims);
        return true;
    }
    bool layerDcmp(const CoordBR &subWorkBR)
    {
        if (!_pMetaData) return false;
        MetaData &m = *_pMetaData;
        m.myrank = 0;
        m.processor_number = 1;
        const int lo = _pNbrhood && !_pNbrhood->empty() ? _pNbrhood->minIRow() : 0;
        const int hi = _pNbrhood && !_pNbrhood->empty() ? _pNbrhood->maxIRow() : 0;
        const int begin = subWorkBR.minIRow() + lo;
        const int end = std::min(subWorkBR.maxIRow(), m._glbDims.nRows() - 1) + hi;
        if (begin < 0 || end >= m._glbDims.nRows() || begin > end) {
            cerr << "RasterLayer::layerDcmp: rows " << begin << ".." << end << " of (" << subWorkBR
                 << ") with the neighborhood halo fall outside the " << m._glbDims.nRows() << " rows of "
                 << _strLayerName << endl;
            return false;
        }
        m._MBR = CoordBR(CellCoord(begin, 0), CellCoord(end, m._glbDims.nCols() - 1));
        m._localdims = SpaceDims(m._MBR.nRows(), m._MBR.nCols());
        if (_pNbrhood && !_pNbrhood->empty()) {
            if (!_pNbrhood->calcWorkBR(m._localworkBR, m._localdims)) return false;
        } else {
            m._localworkBR = CoordBR(CellCoord(0, 0), CellCoord(m._localdims.nRows() - 1, m._localdims.nCols() - 1));
        }
        newCellSpace(m._localdims);
        return true;
    }

    // Copies the metadata of rhs and fills a new cell space with rhs's
    // NoData (rasterLayer.h:740-745).
    bool copyLayerInfo(const RasterLayer<elemType> &rhs)
    {
        if (!copyLayerMetadata(rhs)) return false;
        newCellSpace(_pMetaData->_localdims, (elemType)rhs._pMetaData->noData);
        return true;
    }
    bool copyLayerMetadata(const RasterLayer<elemType> &rhs)
    {
        if (!rhs._pMetaData) return false;
        _meta.reset(new MetaData(*rhs._pMetaData));
        _pMetaData = _meta.get();
        _pMetaData->dataType = dataTypeOf<elemType>();
        if (rhs._pNbrhood) _pNbrhood.reset(new Neighborhood<elemType>(*rhs._pNbrhood));
        return true;
    }
    // Adopts a caller-built MetaData (the reference's operators assign
    // layer._pMetaData = new MetaData() directly).
    void setMetaData(const MetaData &m)
    {
        _meta.reset(new MetaData(m));
        _pMetaData = _meta.get();
    }

    // Writes the layer as a GeoTIFF of the layer's element type, one row per
    // strip, with the input's georeferencing and NoData
    // (rasterLayer.h:958-1087).  The reference returns false without
    // writing for layers read with NON_DCMP; this build writes them too.
    bool writeFile(const char *outputfile)
    {
        if (!_pMetaData || !_pCellSpace) {
            cerr << "[ERROR] RasterLayer::writeFile: layer " << _strLayerName << " has no data" << endl;
            return false;
        }
        if (!syncHost()) return false;
        geotiff::Info info;
        info.rows = _pMetaData->_glbDims.nRows();
        info.cols = _pMetaData->_glbDims.nCols();
        for (int i = 0; i < 6; i++) info.transform[i] = _pMetaData->pTransform[i];
        info.has_nodata = true;
        info.nodata = _pMetaData->noData;
        info.geokeys = _pMetaData->geoKeys.keys;
        info.geodoubles = _pMetaData->geoKeys.doubles;
        info.geoascii = _pMetaData->geoKeys.ascii;
        std::vector<double> values((size_t)_pCellSpace->size());
        const elemType *src = _pCellSpace->data();
        for (size_t i = 0; i < values.size(); i++) values[i] = (double)src[i];
        geotiff::Type t = geotiff::Type::Float64;
        switch (_pMetaData->dataType) {
        case DT_Float32: t = geotiff::Type::Float32; break;
        case DT_Int32: t = geotiff::Type::Int32; break;
        case DT_Byte: t = geotiff::Type::UInt8; break;
        default: break;
        }
        std::string err;
        if (!geotiff::write(outputfile, info, values.data(), t, err)) {
            cerr << "[ERROR] " << err << endl;
            return false;
        }
        if (_pMetaData->geoKeys.empty() && !_pMetaData->projection.empty()) writePamSrs(outputfile);
        return true;
    }

    // Row of `layer` that covers row `row` of this layer, by the ratio of
    // the two layers' row counts, truncated and capped at the other layer's
    // last work row (rasterLayer.h:1289-1319).
    template <class T2>
    int rowAtOtherLayer(RasterLayer<T2> *layer, int row)
    {
        const double scale = (double)layer->metaData()->_glbDims.nRows() / _pMetaData->_glbDims.nRows();
        return (int)std::min(scale * row, (double)layer->metaData()->_localworkBR.maxIRow());
    }
    template <class T2>
    int colAtOtherLayer(RasterLayer<T2> *layer, int col)
    {
        const double scale = (double)layer->metaData()->_glbDims.nCols() / _pMetaData->_glbDims.nCols();
        return (int)std::min(scale * col, (double)layer->metaData()->_localworkBR.maxICol());
    }
    // The reference's test is one-sided: value - noData < 1e-7
    // (rasterLayer.h:33, :347-350).
    bool isNodata(elemType value) const { return (double)value - _pMetaData->noData < 1e-7; }

    MetaData *_pMetaData = nullptr;  ///< public, as in the reference

    // ---- device residency (double layers) --------------------------------
    // The current device copy of the cells, or nullptr.
    const DeviceRaster *deviceCopy() const { return _dev && _dev->current ? _dev.get() : nullptr; }
    // The cells as float32 on the device, uploaded from the host copy unless
    // a current float32 device copy exists; nullptr (with a message) on
    // failure or for a value float32 cannot hold.
    const DeviceRaster *deviceF32() const
    {
        static_assert(std::is_same<elemType, double>::value, "device copies hold double layers");
        if (_dev && _dev->current && _dev->form == DeviceForm::F32 && !_dev->mask) return _dev.get();
        if (!syncHost() || !_pCellSpace || !_pMetaData) return nullptr;
        const SpaceDims &d = _pMetaData->_localdims;
        if (!_dev) _dev.reset(new DeviceRaster());
        if (!_dev->allocate(DeviceForm::F32, d.nRows(), d.nCols(), false)) return nullptr;
        if (!staging::uploadF32(_pCellSpace->data(), (size_t)d.size(), d.nCols(), (float *)_dev->data,
                                _strLayerName.c_str()))
            return nullptr;
        _dev->current = true;
        _dev->newerThanHost = false;
        return _dev.get();
    }
    // Device buffers of `form` (and a NoData mask) for an operator to write
    // the whole layer into; call deviceWritten() once it has.
    DeviceRaster *deviceOutput(DeviceForm form, bool withMask)
    {
        static_assert(std::is_same<elemType, double>::value, "device copies hold double layers");
        if (!_pMetaData) return nullptr;
        const SpaceDims &d = _pMetaData->_localdims;
        if (!_dev) _dev.reset(new DeviceRaster());
        return _dev->allocate(form, d.nRows(), d.nCols(), withMask) ? _dev.get() : nullptr;
    }
    // The device copy now holds the cells; the host copy is stale.
    void deviceWritten()
    {
        if (!_dev) return;
        _dev->current = true;
        _dev->newerThanHost = true;
    }
    // Brings the host CellSpace up to date from a newer device copy.
    bool syncHost() const
    {
        if (!_dev || !_dev->current || !_dev->newerThanHost) return true;
        if constexpr (std::is_same<elemType, double>::value) {
            if (!_pCellSpace || _pCellSpace->nRows() != _dev->rows || _pCellSpace->nCols() != _dev->cols)
                _pCellSpace.reset(new CellSpace<elemType>(SpaceDims(_dev->rows, _dev->cols)));
            if (!staging::download(*_dev, _pMetaData ? _pMetaData->noData : 0.0, _pCellSpace->data())) return false;
            _dev->newerThanHost = false;
            return true;
        } else {
            return false;
        }
    }

  private:
    bool readFileInfo(const char *inputfile, DomDcmpType dcmpType, std::vector<double> *values)
    {
        geotiff::Info info;
        std::string err;
        const bool ok = values ? geotiff::read(inputfile, info, *values, err) : geotiff::read_info(inputfile, info, err);
        if (!ok) {
            cerr << "[ERROR] data file is not open correct: " << err << endl;
            return false;
        }
        _meta.reset(new MetaData());
        _pMetaData = _meta.get();
        MetaData &m = *_pMetaData;
        for (int i = 0; i < 6; i++) m.pTransform[i] = info.transform[i];
        // GDAL reports -1e10 for a band without NoData
        m.noData = info.has_nodata ? info.nodata : -1e10;
        m.row = info.rows;
        m.column = info.cols;
        m.cellSize = (float)m.pTransform[1];
        m.format = "GTiff";
        m._domDcmpType = dcmpType;
        m.dataType = dataTypeOf<elemType>();
        m._glbDims = SpaceDims(info.rows, info.cols);
        m.geoKeys.keys = info.geokeys;
        m.geoKeys.doubles = info.geodoubles;
        m.geoKeys.ascii = info.geoascii;
        return true;
    }
    // GDAL reads a raster's coordinate system from a <file>.aux.xml side
    // file when the GeoTIFF carries no keys.
    void writePamSrs(const char *outputfile)
    {
        std::ofstream os(std::string(outputfile) + ".aux.xml", std::ios::trunc);
        std::string wkt;
        for (char ch : _pMetaData->projection) {
            if (ch == '&') wkt += "&amp;";
            else if (ch == '<') wkt += "&lt;";
            else if (ch == '>') wkt += "&gt;";
            else if (ch == '"') wkt += "&quot;";
            else wkt += ch;
        }
        os << "<PAMDataset>\n  <SRS>" << wkt << "</SRS>\n</PAMDataset>\n";
    }

    // host data replaced or possibly written: the device copy no longer
    // holds the cells (its buffers stay for reuse)
    void retireDevice()
    {
        if (_dev) _dev->current = _dev->newerThanHost = false;
    }

    string _strLayerName;
    int _nLayerID = -1;
    // mutable: bringing the host copy up to date (syncHost) and making the
    // device copy (deviceF32) do not change the layer's cells
    mutable std::unique_ptr<CellSpace<elemType>> _pCellSpace;
    mutable std::unique_ptr<DeviceRaster> _dev;
    std::unique_ptr<Neighborhood<elemType>> _pNbrhood;
    std::unique_ptr<MetaData> _meta;
};

}  // namespace GPRO
