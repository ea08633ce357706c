// pIDW.cpp -- the idw command line (apps/spatial/idw/pIDW.cpp of the
// reference, main :60-343): same flags and the same two decompositions.
//   -dcmp space    (default) the grid is one band and IDW runs at once;
//   -dcmp compute  the load of every granularity x granularity block is
//                  estimated (or read with -readLoad), optionally written
//                  with -writeLoad, and the rows are cut into bands of equal
//                  load; with one process per GPU and one GPU the band is the
//                  whole grid.
// -writeLoad with -dcmp space is the reference's preliminary experiment: the
// output file then holds each cell's cost instead of its value.  The
// reference records CPU milliseconds; the device records the number of
// samples each cell's search examined (its work, deterministic), which
// -dcmp compute -readLoad balances the same way.
#include <new>
#include <cstdlib>
#include <cstring>

#include "idwOperator.h"
#include "idwTransformation.h"

static void Usage(const string &error_msg = "")
{
    if (!error_msg.empty()) cout << "FAILURE: " << error_msg << endl << endl;
    cout << " Usage: idw -sample <sample points file> " << endl
         << "-mask <mask raster file> (optional)" << endl
         << "-out <output raster file>" << endl
         << "-resolution <output resolution (length of one cell)(meter)>" << endl
         << "-fieldIndex <index of the field to be calculated>" << endl
         << "-idwExp <exponent of inverse distance>" << endl
         << "-searchPointNum <maximum of neighbor points to search>" << endl
         << "-searchRange <maximum of range to search>" << endl
         << "-blockSize <size of sample blocks>" << endl
         << "-granularity <compute cells per side of one load cell> (optional, 10)" << endl
         << "-dcmp <space|compute>" << endl
         << "-writeLoad <path to write load file>" << endl
         << "-readLoad <path to read load file>" << endl
         << endl;
    exit(1);
}

int main(int argc, char *argv[])
try {
    const char *inputFileName = nullptr, *maskFileName = nullptr, *outputFileName = nullptr;
    const char *writeLoadPath = nullptr, *readLoadPath = nullptr;
    float cellSize = 0;
    int fldIdx = -1, idw_nbrPoints = 0, idw_power = 0, granularity = 10;
    double idw_buffer = 0, blockSize = 0;
    bool decomposeBySpace = true;
    bool have[6] = {false, false, false, false, false, false};
    for (int i = 1; i < argc; i += 2) {
        const string a = argv[i];
        if (i + 1 >= argc) Usage("No argument followed '" + a + "'!");
        const char *v = argv[i + 1];
        if (a == "-sample") inputFileName = v;
        else if (a == "-mask") maskFileName = v;
        else if (a == "-out") outputFileName = v;
        else if (a == "-resolution") cellSize = (float)atof(v), have[0] = true;
        else if (a == "-fieldIndex") fldIdx = atoi(v), have[1] = true;
        else if (a == "-idwExp") idw_power = atoi(v), have[2] = true;
        else if (a == "-searchPointNum") idw_nbrPoints = atoi(v), have[3] = true;
        else if (a == "-searchRange") idw_buffer = atof(v), have[4] = true;
        else if (a == "-blockSize") blockSize = atof(v), have[5] = true;
        else if (a == "-granularity") granularity = atoi(v);
        else if (a == "-dcmp") decomposeBySpace = ifDecomposeBySpace(v);
        else if (a == "-writeLoad") writeLoadPath = v;
        else if (a == "-readLoad") readLoadPath = v;
        else Usage("unknown option '" + a + "'");
    }
    if (!inputFileName || !outputFileName) Usage("-sample and -out are required");
    for (bool h : have)
        if (!h) Usage("-resolution, -fieldIndex, -idwExp, -searchPointNum, -searchRange and -blockSize are required");
    if (!Application::START(HIP_Type, argc, argv)) return 1;

    double starttime;
    IDWOperator idwOper(cellSize, idw_nbrPoints, idw_power, idw_buffer, blockSize);
    char *spatialrefWkt = nullptr;
    idwOper.readSampleNums(inputFileName, &spatialrefWkt);
    vector<SamplePoint> samples;
    if (!idwOper.readSamples(inputFileName, fldIdx, &spatialrefWkt, samples)) return 1;
    idwOper.creatSampleBlocks(samples);
    if (!idwOper.plan()) return 1;
    vector<SamplePoint>().swap(samples);
    RasterLayer<int> maskLayer("maskLayer");
    maskLayer.newLocalNbrhood();
    RasterLayer<double> idwLayer("idwLayer");
    idwLayer.newLocalNbrhood();
    if (decomposeBySpace) {
        idwOper.idwLayer(idwLayer, &spatialrefWkt);
        if (maskFileName) {
            if (!maskLayer.readFile(maskFileName, ROWWISE_DCMP)) return 1;
            idwOper.maskLayer(maskLayer);
        }
        // preliminary experiment: the output layer records each cell's cost
        if (writeLoadPath) idwOper._writePreExpLoad = true;
        cout << "start computing" << endl;
        starttime = Application::wtime();
        if (!idwOper.Run()) return 1;
    } else {
        idwOper.idwLayerSerial(idwLayer, &spatialrefWkt);
        CoordBR subWorkBR;
        cout << "start dcmp" << endl;
        starttime = Application::wtime();
        ComputeLayer<double> comptLayer("computLayer");
        comptLayer.addRasterLayerSerial(&idwLayer);
        if (!comptLayer.init(nullptr, granularity)) return 1;
        RasterLayer<int> fullMaskLayer("fullMaskLayer");
        if (readLoadPath) {
            if (!comptLayer.readComputeLoadFile(readLoadPath)) return 1;
        } else {
            if (maskFileName) {
                fullMaskLayer.newLocalNbrhood();
                if (!fullMaskLayer.readGlobalFileSerial(maskFileName)) return 1;
                idwOper.maskLayer(fullMaskLayer);
            }
            IdwTransformation idwTrans(&comptLayer, &idwOper);
            if (!idwTrans.run()) return 1;
        }
        if (!comptLayer.getCompuLoad(ROWWISE_DCMP, 1, subWorkBR)) return 1;
        if (!readLoadPath && writeLoadPath && !comptLayer.writeComputeIntensityFile(writeLoadPath)) return 1;
        cout << "dcmp time is " << Application::wtime() - starttime << endl;
        cout << 0 << " subWorkBR " << subWorkBR.minIRow() << " " << subWorkBR.maxIRow() << " " << subWorkBR.nRows()
             << endl;
        idwOper.idwLayer(idwLayer, &spatialrefWkt, subWorkBR);
        if (maskFileName) {
            if (!maskLayer.readFile(maskFileName, subWorkBR, ROWWISE_DCMP)) return 1;
            idwOper.maskLayer(maskLayer);
        }
        cout << "start computing" << endl;
        starttime = Application::wtime();
        if (!idwOper.Run()) return 1;
    }
    cout << "compute time is " << Application::wtime() - starttime << endl << endl;
    if (!idwLayer.writeFile(outputFileName)) return 1;
    cout << "write done." << endl;
    return Application::END() ? 0 : 1;
} catch (const std::bad_alloc &) {
    // the reference stops with a message where an allocation fails
    cerr << "[ERROR] out of host memory" << endl;
    return 1;
}
