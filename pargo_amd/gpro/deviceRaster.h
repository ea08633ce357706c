// deviceRaster.h -- the device-resident copy of a RasterLayer's cells.
//
// The reference's CellSpace<T> is one host matrix per rank
// (gpro/cellSpace.h:117-234) that every operator reads and writes.  Here a
// layer can also hold its cells in HBM in the form the kernels use:
//   F32     float32 values (DEMs; the slope output with a NoData bitmask),
//   CODES8  uint8 D8 codes + NoData bitmask (the D8 output; dirType 8 or 256),
//   F64     float64 values (contributing area, filled DEM).
// A device operator takes its input from the device copy when it is
// current and leaves its output there, marking the host copy stale; the
// host copy is rebuilt (one copy back, converted on the host's cores) only
// when host code asks for the cells.  So operators chained through the API
// (DEM -> D8 -> SCA, DEM -> slope) never round-trip through the host.
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/pargo.h"
#include "application.h"

namespace GPRO {

enum class DeviceForm { None, F32, Codes8, F64 };

// Bytes moved between host and device by the gpro layer (all layers), for
// tests and the command lines' reports.
struct DeviceTransfers {
    static unsigned long long h2d, d2h;
};

class DeviceRaster {
  public:
    DeviceRaster() {}
    ~DeviceRaster() { release(); }
    DeviceRaster(const DeviceRaster &) = delete;
    DeviceRaster &operator=(const DeviceRaster &) = delete;

    // Allocates rows x cols cells of `form` (plus a NoData bitmask when
    // withMask), reusing the buffers when the shape and form are unchanged.
    // The contents are not current until an operator writes them.
    bool allocate(DeviceForm f, int r, int c, bool withMask);
    // Frees the buffers (after the work queued on them has finished).
    void release();
    // Frees every device raster of the process; Application::END calls it
    // before it closes the context.
    static void releaseAll();

    static size_t elemBytes(DeviceForm f) { return f == DeviceForm::F32 ? 4 : (f == DeviceForm::Codes8 ? 1 : 8); }
    static size_t maskWords(int c) { return (size_t)(c + 31) / 32; }

    DeviceForm form = DeviceForm::None;
    void *data = nullptr;
    uint32_t *mask = nullptr;  // bit c % 32 of word c / 32 of a row: NoData
    int rows = 0, cols = 0;
    int dirType = 256;         // CODES8: 8 (window index, 255 flat) or 256 (ESRI bits, 0 flat)
    bool current = false;      // the device copy holds the layer's cells
    bool newerThanHost = false;  // the host copy is stale
};

// Host <-> device staging through pinned buffers in chunks, with the
// element conversion spread over the host's cores.  Defined in
// deviceRaster.cpp.
namespace staging {
// float64 host cells -> float32 device cells; refuses (message, false) a
// value float32 cannot hold, naming its row and column.
bool uploadF32(const double *src, size_t n, int cols, float *dst, const char *what);
// float64 host cells -> float64 device cells.
bool uploadF64(const double *src, size_t n, double *dst);
// device cells of `form` (+ mask) -> float64 host cells: NoData cells (mask
// bit set) get `nodata`; CODES8 flat codes get -1.
bool download(const DeviceRaster &d, double nodata, double *dst);
}  // namespace staging

}  // namespace GPRO
