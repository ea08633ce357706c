// pSCA.cpp -- the scad8 command line (apps/hydrology/scad8/pSCA.cpp of the
// reference, main :34-82).
//
// scad8 <D8 file> <neighbor file> <output file> as in the reference (three
// positional arguments, otherwise "please input right parameter." and exit
// status 0).  The D8 file holds window-index codes (flowdird8 with
// dirType 8).  A fourth argument "-converge" selects the converged
// accumulation instead of the reference's single sweep.  The output is
// written as a Float64 GeoTIFF with the D8 layer's georeferencing and
// NoData; the reference reads the D8 file without a decomposition type, for
// which its writeFile returns without writing (gpro/rasterLayer.h:1009-1012).
#include <new>
#include <cstring>

#include "scaOperator.h"

int main(int argc, char *argv[])
try {
    bool converge = false;
    if (argc == 5 && !strcmp(argv[4], "-converge")) {
        converge = true;
        argc = 4;
    }
    if (argc != 4) {
        cerr << "please input right parameter.";
        return 0;
    }
    const char *inputfilename = argv[1], *neighborfile = argv[2], *outputfilename = argv[3];
    if (!Application::START(HIP_Type, argc, argv)) return 1;

    RasterLayer<double> d8Layer("d8Layer");
    if (!d8Layer.readNeighborhood(neighborfile)) return 1;
    if (!d8Layer.readFile(inputfilename)) return 1;
    RasterLayer<double> scaLayer("scaLayer");
    scaLayer.copyLayerInfo(d8Layer);

    const double starttime = Application::wtime();
    SCAOperator scaOper;
    scaOper.setConverge(converge);
    scaOper.d8Layer(d8Layer);
    scaOper.scaLayer(scaLayer);
    if (!scaOper.Run()) return 1;
    const double endtime = Application::wtime();
    cout << "run time is " << endtime - starttime << endl;

    if (!scaLayer.writeFile(outputfilename)) return 1;
    return Application::END() ? 0 : 1;
} catch (const std::bad_alloc &) {
    // the reference stops with a message where an allocation fails
    cerr << "[ERROR] out of host memory" << endl;
    return 1;
}
