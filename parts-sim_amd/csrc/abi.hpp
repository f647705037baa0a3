// Shared pieces of the C ABI implementation: the thread-local last error, the exception to status
// mapping every entry point runs under, and option defaults.
#pragma once

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/psx.h"
#include "model.hpp"

// the opaque handle of include/psx.h: one device-resident model
struct psx_model_s {
    std::unique_ptr<psx::Model> m;
};

namespace psx::abi {
inline thread_local std::string g_error;

inline int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

template <class F> inline int guarded(F&& fn) {
    try {
        return fn();
    } catch (const psx::NoDevice& e) {
        return fail(PSX_ERR_NO_DEVICE, e.what());
    } catch (const psx::DeviceError& e) {
        return fail(PSX_ERR_DEVICE, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(PSX_ERR_ARGUMENT, e.what());
    } catch (const std::runtime_error& e) {
        const std::string w = e.what();
        return fail(w.find("direct solver") != std::string::npos ? PSX_ERR_NO_COARSE : PSX_ERR_INTERNAL, w);
    } catch (const std::exception& e) {
        return fail(PSX_ERR_INTERNAL, e.what());
    } catch (...) {
        return fail(PSX_ERR_INTERNAL, "unknown error");
    }
}

inline psx_solve_opts defaults(const psx_solve_opts* o) {
    psx_solve_opts d{};
    d.tol = 1e-6;
    d.max_iter = 500;
    d.precision = PSX_PREC_FP64;
    if (o) {
        d = *o;
        if (!(d.tol > 0)) d.tol = 1e-6;
        if (d.precision < PSX_PREC_FP64 || d.precision > PSX_PREC_MIXED) throw std::invalid_argument("unknown precision");
    }
    return d;
}

inline void checkFinite(const double* v, int64_t n, const char* what) {
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(v[i])) throw std::invalid_argument(std::string(what) + " must be finite numbers.");
}
}  // namespace psx::abi
