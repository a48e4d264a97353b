// Internal helpers of the ket:: host layer (not installed).
#pragma once

#include <optional>
#include <random>
#include <stdexcept>
#include <string>

#include "kq.h"
#include "kettle/state/statevector.hpp"

namespace ket::internal
{

// libkq status -> kettle's exception types: out-of-range stays std::out_of_range, everything
// else becomes std::runtime_error with libkq's message (which is kettle's text for kettle errors)
inline void kq_check(int status)
{
    if (status == KQ_OK) {
        return;
    }
    const std::string msg = kq_last_error();
    if (status == KQ_ERR_OUT_OF_RANGE) {
        throw std::out_of_range {msg};
    }
    throw std::runtime_error {msg};
}

// get_prng_ (common/prng.cpp:9-19)
inline auto get_prng(std::optional<int> seed) -> std::mt19937
{
    if (seed) {
        return std::mt19937 {static_cast<std::mt19937::result_type>(seed.value())};
    }
    auto device = std::random_device {};
    return std::mt19937 {device()};
}

inline auto kq_precision(Precision p) -> int { return p == Precision::C64 ? KQ_C64 : KQ_C128; }

struct KqStateDeleter
{
    void operator()(kq_state* s) const noexcept { (void) kq_state_destroy(s); }
};

}  // namespace ket::internal
