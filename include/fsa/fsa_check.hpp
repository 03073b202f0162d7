// fsa_check.hpp -- turns a non-zero libfsa status into the reference's
// behaviour on bad input (an assert-style abort), printing the library message.
#ifndef FSA_CHECK_HPP
#define FSA_CHECK_HPP

#include <cstdio>
#include <cstdlib>

#include "../fsa.h"

static inline void fsa_check(int status, const char* what) {
    if (status != FSA_OK) {
        std::fprintf(stderr, "%s failed: %s\n", what, fsa_last_error());
        std::abort();
    }
}

#endif  // FSA_CHECK_HPP
