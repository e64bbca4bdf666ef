// Shared helpers of the C ABI translation units.
#pragma once

#include <new>
#include <string>

#include "../../../include/mdft.h"
#include "runtime.hpp"

namespace mdft {
void set_error(const std::string& s);
const char* get_error();
}  // namespace mdft

#define MDFT_API_BEGIN try {
#define MDFT_API_END                                   \
    }                                                  \
    catch (const mdft::Error& e)                       \
    {                                                  \
        mdft::set_error(e.what());                     \
        return e.code();                               \
    }                                                  \
    catch (const std::bad_alloc&)                      \
    {                                                  \
        mdft::set_error("host allocation failed");     \
        return MDFT_ERR_INVALID;                       \
    }                                                  \
    catch (const std::exception& e)                   \
    {                                                  \
        mdft::set_error(e.what());                     \
        return MDFT_ERR_INVALID;                       \
    }                                                  \
    return MDFT_OK;
