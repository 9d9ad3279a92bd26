/* Filling a prr_error: shared by the host sources of proresraw_amd. */
#ifndef PRR_FAIL_H
#define PRR_FAIL_H

#include <cstdint>
#include <cstring>

#include "../../include/proresraw_amd.h"

/* Sets *err (when given) to `code` with its payload and returns `code`. */
static inline int fail(prr_error *err, int code, const char *what = nullptr, uint64_t a = 0,
                       uint64_t b = 0)
{
    if (err) {
        std::memset(err, 0, sizeof *err);
        err->code = code;
        if (what)
            std::strncpy(err->what, what, sizeof err->what - 1);
        err->a = a;
        err->b = b;
    }
    return code;
}

#endif
