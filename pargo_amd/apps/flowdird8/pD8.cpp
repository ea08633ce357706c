// pD8.cpp -- the flowdird8 command line (apps/hydrology/flowdird8/pD8.cpp
// of the reference, main :47-139).
//
// Simple usage as in the reference: flowdird8 <dem> <nbr> <out> <dirType>.
// The full usage takes -elev, -nbr, -out and the -dirtype flag that the
// reference's usage text names (its parser never reads it).  The output is
// written as a Float64 GeoTIFF; the reference reads the DEM without a
// decomposition type, for which its writeFile returns without writing
// (gpro/rasterLayer.h:1009-1012).
#include <new>
#include <cstdlib>
#include <cstring>

#include "d8Operator.h"

static void Usage(const string &error_msg = "")
{
    if (!error_msg.empty()) cout << "FAILURE: " << error_msg << endl << endl;
    cout << " Usage: flowdird8 -elev <elevation grid file> -nbr <neighbor definition file> -out <output flow "
            "direction file> -dirtype <direction type>"
         << endl;
    cout << " direction type is the representation type of flow direction result, should be one of the following:"
         << endl;
    cout << " \t 8: the median of the tied steepest window indices 0..8 is output." << endl;
    cout << " \t 256: If the maximum descent to several cells is the same, the direction of neighborhood will be "
            "accumulated."
         << endl;
    cout << " Or use the Simple Usage: flowdird8 <elevation grid file> <neighbor definition file> <output flow "
            "direction file> <direction type>"
         << endl;
    exit(1);
}

int main(int argc, char *argv[])
try {
    if (argc < 5) Usage("Too few arguments to run this program.");
    const char *inputfilename = nullptr, *neighborfile = nullptr, *outputfilename = nullptr;
    int dirType = 0;
    bool simpleusage = true;
    for (int i = 1; i < argc;) {
        const char *a = argv[i];
        const char **target = nullptr;
        if (!strcmp(a, "-elev")) target = &inputfilename;
        else if (!strcmp(a, "-nbr")) target = &neighborfile;
        else if (!strcmp(a, "-out")) target = &outputfilename;
        if (target || !strcmp(a, "-dirtype")) {
            simpleusage = false;
            if (i + 1 >= argc) Usage(string("No argument followed '") + a + "'!");
            if (target) *target = argv[i + 1];
            else dirType = atoi(argv[i + 1]);
            i += 2;
            continue;
        }
        if (!simpleusage) Usage("DO NOT mix the Full and Simple usages!");
        inputfilename = argv[1];
        neighborfile = argv[2];
        outputfilename = argv[3];
        dirType = atoi(argv[4]);
        break;
    }
    if (!inputfilename || !FileExists(inputfilename)) Usage("The input DEM file not exists");
    if (!neighborfile || !FileExists(neighborfile)) Usage("neighbor file not exists");
    if (!outputfilename) Usage("No output flow direction file given");

    if (!Application::START(HIP_Type, argc, argv)) return 1;
    if (!(dirType == 8 || dirType == 256)) {
        cerr << "please input right parameter.";
        Application::END();
        return 0;
    }
    RasterLayer<double> demLayer("demLayer");
    if (!demLayer.readNeighborhood(neighborfile)) return 1;
    if (!demLayer.readFile(inputfilename)) return 1;
    RasterLayer<double> sflowLayer("d8Layer");
    sflowLayer.copyLayerInfo(demLayer);

    const double starttime = Application::wtime();
    D8Operator SFlOper(dirType);
    SFlOper.demLayer(demLayer);
    SFlOper.d8Layer(sflowLayer);
    if (!SFlOper.Run()) return 1;
    const double endtime = Application::wtime();
    cout << "run time is " << endtime - starttime << endl;
    if (!sflowLayer.writeFile(outputfilename)) return 1;
    return Application::END() ? 0 : 1;
} catch (const std::bad_alloc &) {
    // the reference stops with a message where an allocation fails
    cerr << "[ERROR] out of host memory" << endl;
    return 1;
}
