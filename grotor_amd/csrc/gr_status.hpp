// Status codes and error text for the C ABI (the reference throws
// std::runtime_error and main maps it to exit 1, include/cuda_memory.hpp:12-16,
// src/main.cu:168-178; here the exception becomes a status code).
#pragma once

#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/grotor_amd.h"

namespace grotor {

struct GrError : std::runtime_error {
    int code;
    GrError(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

void set_last_error(const std::string& message);

template <class F>
int guarded(F&& f) {
    try {
        f();
        set_last_error("");
        return GR_OK;
    } catch (const GrError& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return GR_E_ARG;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return GR_E_CAPACITY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return GR_E_IO;
    }
}

}  // namespace grotor
