// application.h -- program start and end of the gpro interface
// (gpro/application.h:30-45, application.cpp:5-33 of the reference).
//
// START records the program type and opens the GPU context that the
// device operators share; END waits for outstanding work and closes it.
// There is no MPI here: this build runs one process per GPU and one GPU.
#pragma once

#include "basicTypes.h"

struct pargo_ctx;

namespace GPRO {

class Application {
  public:
    static bool START(ProgramType programType, int argc, char *argv[]);
    static bool END();
    // The context opened by START (opened on first use when START was not
    // called).  Returns nullptr and prints the reason if the GPU library
    // cannot open a device.
    static pargo_ctx *device();
    // The open context, or nullptr (never opens one).
    static pargo_ctx *current() { return _ctx; }
    // Wall-clock seconds (the reference's MPI_Wtime).
    static double wtime();

    static ProgramType _programType;

  private:
    static pargo_ctx *_ctx;
};

// Prints the library's message for a failed call and returns false;
// returns true for PARGO_OK.
bool deviceOk(int status, const char *what);

// Common helpers of the reference's apps (gpro/utility.h).
bool FileExists(const char *path);
bool StringMatch(const string &a, const string &b);  // case-insensitive

}  // namespace GPRO
