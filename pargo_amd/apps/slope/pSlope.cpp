// pSlope.cpp -- the slope command line (apps/morphology/slope/pSlope.cpp of
// the reference, main :52-154): same flags, same simple usage, same output
// (tan of the slope angle as a Float64 GeoTIFF with the DEM's NoData).
#include <new>
#include <cstring>

#include "slopeOperator.h"

static void Usage(const string &error_msg = "")
{
    if (!error_msg.empty()) cout << "FAILURE: " << error_msg << endl << endl;
    cout << " Usage: slope -elev <elevation grid file> -nbr <neighbor definition file> "
            "-slp <output slope file> [-mtd <algorithm>]"
         << endl;
    cout << "The available algorithm for slope are:" << endl;
    cout << "\t FD: (default) Third-order finite difference weighted by reciprocal of squared distance" << endl;
    cout << "\t FFD: Frame finite difference" << endl;
    cout << "\t MD: Maximum downslope" << endl;
    cout << "\t SD: Simple difference" << endl;
    cout << "\t SFD: Second-order finite difference" << endl;
    cout << "\t TFD: Third-order finite difference" << endl;
    cout << "\t TFDW: Third-order finite difference weighted by reciprocal of distance" << endl << endl;
    cout << " Or use the Simple Usage: slope <elevation grid file> <neighbor definition file> "
            "<output slope file> [<algorithm>]"
         << endl;
    exit(1);
}

int main(int argc, char *argv[])
try {
    if (argc < 4) Usage("Too few arguments to run this program.");
    const char *inputfilename = nullptr, *neighborfile = nullptr, *outputfilename = nullptr;
    SlopeAlgor calcalgor = FD;
    bool simpleusage = true;
    for (int i = 1; i < argc;) {
        const char *a = argv[i];
        const char **target = nullptr;
        if (!strcmp(a, "-elev")) target = &inputfilename;
        else if (!strcmp(a, "-nbr")) target = &neighborfile;
        else if (!strcmp(a, "-slp")) target = &outputfilename;
        if (target || !strcmp(a, "-mtd")) {
            simpleusage = false;
            if (i + 1 >= argc) Usage(string("No argument followed '") + a + "'!");
            if (target) *target = argv[i + 1];
            else calcalgor = GetSlopeAlgorithm(argv[i + 1]);
            i += 2;
            continue;
        }
        if (!simpleusage) Usage("DO NOT mix the Full and Simple usages!");
        inputfilename = argv[1];
        neighborfile = argv[2];
        outputfilename = argv[3];
        if (argc >= 5) calcalgor = GetSlopeAlgorithm(argv[4]);
        break;
    }
    if (!inputfilename || !FileExists(inputfilename)) Usage("The input DEM file not exists");
    if (!neighborfile || !FileExists(neighborfile)) Usage("neighbor file not exists");
    if (!outputfilename) Usage("No output slope file given");

    if (!Application::START(HIP_Type, argc, argv)) return 1;
    RasterLayer<double> demLayer("demLayer");
    if (!demLayer.readNeighborhood(neighborfile)) return 1;
    if (!demLayer.readFile(inputfilename, ROWWISE_DCMP)) return 1;
    RasterLayer<double> slopeLayer("slopeLayer");
    slopeLayer.copyLayerInfo(demLayer);

    const double starttime = Application::wtime();
    SlopeOperator slpOper;
    slpOper.demLayer(demLayer);
    slpOper.slopeLayer(slopeLayer);
    slpOper.calcAlgorithm(calcalgor);
    if (!slpOper.Run()) return 1;
    const double endtime = Application::wtime();
    cout << "run time is " << endtime - starttime << endl;
    if (!slopeLayer.writeFile(outputfilename)) return 1;
    return Application::END() ? 0 : 1;
} catch (const std::bad_alloc &) {
    // the reference stops with a message where an allocation fails
    cerr << "[ERROR] out of host memory" << endl;
    return 1;
}
