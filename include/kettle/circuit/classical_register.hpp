#pragma once

#include <cstddef>
#include <optional>
#include <vector>

// Measured bits of a simulation (include/kettle/circuit/classical_register.hpp in the reference).
namespace ket
{

class ClassicalRegister
{
public:
    explicit ClassicalRegister(std::size_t n_bits) : measured_bits_(n_bits, std::nullopt) {}

    [[nodiscard]] auto is_measured(std::size_t index) const -> bool { return measured_bits_[index].has_value(); }
    [[nodiscard]] auto get(std::size_t index) const -> int { return measured_bits_[index].value(); }
    void set(std::size_t index, int value) { measured_bits_[index] = value; }
    [[nodiscard]] auto size() const -> std::size_t { return measured_bits_.size(); }

private:
    std::vector<std::optional<int>> measured_bits_;
};

}  // namespace ket
