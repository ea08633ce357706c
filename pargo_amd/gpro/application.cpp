// application.cpp -- see application.h.
#include "application.h"

#include <sys/stat.h>

#include <cctype>
#include <chrono>

#include "../../include/pargo.h"
#include "deviceRaster.h"

namespace GPRO {

ProgramType Application::_programType = MPI_Type;
pargo_ctx *Application::_ctx = nullptr;

bool Application::START(ProgramType programType, int, char *[])
{
    _programType = programType;
    return device() != nullptr;
}

bool Application::END()
{
    if (!_ctx) return true;
    const bool ok = deviceOk(pargo_sync(_ctx), "pargo_sync");
    DeviceRaster::releaseAll();  // device copies of layers that outlive END
    pargo_destroy(_ctx);
    _ctx = nullptr;
    return ok;
}

pargo_ctx *Application::device()
{
    if (_ctx) return _ctx;
    if (pargo_abi_version() != PARGO_ABI_VERSION) {
        cerr << "[ERROR] libpargo ABI " << pargo_abi_version() << " does not match the header's " << PARGO_ABI_VERSION
             << endl;
        return nullptr;
    }
    if (pargo_create(0, &_ctx) != PARGO_OK) {
        cerr << "[ERROR] cannot open the GPU: " << pargo_last_error(nullptr) << endl;
        _ctx = nullptr;
    }
    return _ctx;
}

double Application::wtime()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

bool deviceOk(int status, const char *what)
{
    if (status == PARGO_OK) return true;
    const char *msg = pargo_last_error(Application::device());
    cerr << "[ERROR] " << what << " failed (" << status << "): " << (msg && *msg ? msg : "no message") << endl;
    return false;
}

bool FileExists(const char *path)
{
    struct stat st;
    return path && stat(path, &st) == 0 && S_ISREG(st.st_mode);
}

bool StringMatch(const string &a, const string &b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}

}  // namespace GPRO
