// pPitRemove.cpp -- the pitremove command line (apps/hydrology/pitremove/
// pPitRemove.cpp of the reference, Usage :30-41, main :43-133).
//
//   pitremove -elev <dem> -nbr <nbr> -out <filled>
//   pitremove <dem> <nbr> <filled>
// The filled DEM is written as a Float64 GeoTIFF with the DEM's
// georeferencing and NoData.
#include <new>
#include <cstdlib>
#include <cstring>

#include "pitRemoveOperator.h"

static void Usage(const string &error_msg = "")
{
    if (!error_msg.empty()) cout << "FAILURE: " << error_msg << endl << endl;
    cout << " Usage: pitremove -elev <elevation grid file> -nbr <neighbor definition file> -out <pit-removed output "
            "file>"
         << endl;
    cout << " Or use the Simple Usage: pitremove <elevation grid file> <neighbor definition file> <pit-removed "
            "output file>"
         << endl
         << endl;
    cout << "Example.1. pitremove -elev /path/to/elev.tif -nbr /path/to/moore.nbr -out /path/to/pitRemoved.tif"
         << endl;
    cout << "Example.2. pitremove /path/to/elev.tif /path/to/moore.nbr /path/to/pitRemoved.tif" << endl;
    exit(1);
}

int main(int argc, char *argv[])
try {
    if (argc < 4) Usage("Too few arguments to run this program.");
    const char *inputfilename = nullptr, *neighborfile = nullptr, *outputfilename = nullptr;
    bool simpleusage = true;
    for (int i = 1; i < argc;) {
        const char *a = argv[i];
        const char **target = nullptr;
        if (!strcmp(a, "-elev")) target = &inputfilename;
        else if (!strcmp(a, "-nbr")) target = &neighborfile;
        else if (!strcmp(a, "-out")) target = &outputfilename;
        if (target) {
            simpleusage = false;
            if (i + 1 >= argc) Usage(string("No argument followed '") + a + "'!");
            *target = argv[i + 1];
            i += 2;
            continue;
        }
        if (!simpleusage) Usage("DO NOT mix the Full and Simple usages!");
        inputfilename = argv[1];
        neighborfile = argv[2];
        outputfilename = argv[3];
        break;
    }
    if (!inputfilename || !FileExists(inputfilename)) Usage("The input DEM file not exists");
    if (!neighborfile || !FileExists(neighborfile)) Usage("neighbor file not exists");
    if (!outputfilename) Usage("No output file given");

    if (!Application::START(HIP_Type, argc, argv)) return 1;
    RasterLayer<double> demLayer("demLayer");
    if (!demLayer.readNeighborhood(neighborfile)) return 1;
    if (!demLayer.readFile(inputfilename, ROWWISE_DCMP)) return 1;
    RasterLayer<double> wdemLayer("wdemLayer");
    if (!wdemLayer.readNeighborhood(neighborfile)) return 1;
    wdemLayer.copyLayerInfo(demLayer);

    const double starttime1 = Application::wtime();
    PitRemoveOperator PitOper;
    PitOper.demLayer(demLayer);
    PitOper.wdemLayer(wdemLayer);
    if (!PitOper.Run()) return 1;
    const double endtime1 = Application::wtime();
    cout << "run time is " << endtime1 - starttime1 << endl;

    if (!wdemLayer.writeFile(outputfilename)) return 1;
    return Application::END() ? 0 : 1;
} catch (const std::bad_alloc &) {
    // the reference stops with a message where an allocation fails
    cerr << "[ERROR] out of host memory" << endl;
    return 1;
}
