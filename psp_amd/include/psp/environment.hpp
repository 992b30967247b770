// Environment: psp::Initialize/Finalize (src/global.cpp:24-53), the error
// types (LogicError/RuntimeError) and a single-process communicator.
#pragma once

#include <complex>
#include <sstream>
#include <stdexcept>
#include <string>

#include "psp_hip.h"

namespace psp {

template <typename Real>
using Complex = std::complex<Real>;

namespace mpi {
// One GPU holds the whole grid: the communicator is a single rank.
struct Comm {
    int rank = 0, size = 1;
};
inline const Comm COMM_WORLD{};
inline int CommRank(const Comm& c) { return c.rank; }
inline int CommSize(const Comm& c) { return c.size; }
}  // namespace mpi

inline void Initialize(int& /*argc*/, char**& /*argv*/) {}
inline void Finalize() {}

template <typename... Args>
[[noreturn]] inline void LogicError(Args&&... args)
{
    std::ostringstream os;
    (os << ... << args);
    throw std::logic_error(os.str());
}

template <typename... Args>
[[noreturn]] inline void RuntimeError(Args&&... args)
{
    std::ostringstream os;
    (os << ... << args);
    throw std::runtime_error(os.str());
}

// Maps a C-ABI status to the reference's exception types.
inline void ThrowOnStatus(int status, const psp_hip_ctx* ctx, const char* what)
{
    if (status == PSP_HIP_OK) return;
    const std::string msg = std::string(what) + ": " + psp_hip_last_error(ctx);
    if (status == PSP_HIP_ERR_INVALID || status == PSP_HIP_ERR_SHALLOW ||
        status == PSP_HIP_ERR_STATE)
        throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

}  // namespace psp
