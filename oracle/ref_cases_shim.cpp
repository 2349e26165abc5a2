// SPDX-License-Identifier: MIT
// Test-only bridge to the reference's portable AVBD fixtures.  Compiled by
// oracle/Makefile against tests/solver/geometry_constraint_cases.hpp where
// it lies in the reference checkout; the resulting library goes to
// oracle/_ref/ and is never part of the product.  It exposes the fixture
// inputs (make_block), the reference mirror's outputs (evaluate) and the
// reference's own acceptance check (validate) so the tests can judge the
// oracle with the reference's expected values and independent double
// solve (geometry_constraint_cases.hpp:484-641).
#include "geometry_constraint_cases.hpp"

#include <cstring>

extern "C" {

unsigned ref_case_count() { return avbd_cases::count; }
unsigned ref_block_count() { return avbd_cases::block_count; }

const char *ref_case_name(unsigned index) { return avbd_cases::name(index); }

// Packed lower triangle (row*(row+1)/2+col) and gradient of block fixture.
void ref_make_block(unsigned index, float h[21], float g[6]) {
    const auto block = avbd_cases::make_block(index);
    for (unsigned r = 0; r < 6; ++r) {
        g[r] = block.g[r];
        for (unsigned c = 0; c <= r; ++c) h[r * (r + 1) / 2 + c] = block.h[r][c];
    }
}

int ref_evaluate(unsigned index, float values[8]) {
    const auto result = avbd_cases::evaluate(index);
    std::memcpy(values, result.value, sizeof(result.value));
    return result.solved ? 1 : 0;
}

int ref_validate(unsigned index, const float values[8], int solved) {
    avbd_cases::Result result{};
    std::memcpy(result.value, values, sizeof(result.value));
    result.solved = solved != 0;
    return avbd_cases::validate(index, result) ? 1 : 0;
}

}
