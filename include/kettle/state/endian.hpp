#pragma once

#include <cstdint>

namespace ket
{

enum class Endian : std::uint8_t
{
    LITTLE,
    BIG
};

}  // namespace ket
