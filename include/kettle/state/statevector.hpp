#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "kettle/common/mathtools.hpp"
#include "kettle/common/tolerance.hpp"
#include "kettle/state/endian.hpp"
#include "kettle/state/qubit_state_conversion.hpp"

struct kq_state;

/*
    kettle's Statevector (include/kettle/state/statevector.hpp:14-107 in the reference), resident
    in device memory.  The amplitudes live on the GPU while circuits run; a host copy is made on
    demand (operator[], at(), copies) and uploaded again only after host-side writes.  A state
    created as |0...0> or from a bitstring is materialised directly on whichever side needs it.
    Precision::C64 stores complex<float> amplitudes; reads still return complex<double>.
*/
namespace ket
{

enum class Precision : std::uint8_t
{
    C128,
    C64
};

class Statevector
{
public:
    explicit Statevector(std::size_t n_qubits, Precision precision = Precision::C128);
    explicit Statevector(std::vector<std::complex<double>> coefficients, Endian input_endian = Endian::LITTLE,
                         double normalization_tolerance = CONSTRUCTION_NORMALIZATION_TOLERANCE);
    explicit Statevector(const std::string& computational_state, Endian input_endian = Endian::LITTLE);

    Statevector(const Statevector& other);
    Statevector(Statevector&& other) noexcept;
    auto operator=(const Statevector& other) -> Statevector&;
    auto operator=(Statevector&& other) noexcept -> Statevector&;
    ~Statevector();

    auto operator[](std::size_t index) const -> const std::complex<double>&;
    auto operator[](std::size_t index) -> std::complex<double>&;
    [[nodiscard]] auto at(std::size_t index) const -> const std::complex<double>&;
    auto at(std::size_t index) -> std::complex<double>&;
    [[nodiscard]] auto at(const std::string& bitstring, Endian endian = Endian::LITTLE) const -> const std::complex<double>&;

    [[nodiscard]] auto n_states() const noexcept -> std::size_t { return n_states_; }
    [[nodiscard]] auto n_qubits() const noexcept -> std::size_t { return n_qubits_; }
    [[nodiscard]] auto precision() const noexcept -> Precision { return precision_; }

    // all amplitudes on the host (downloads if the device copy is newer)
    [[nodiscard]] auto coefficients() const -> const std::vector<std::complex<double>>&;

    // engine access: the device state, up to date; after device-side writes call mark_device_modified()
    [[nodiscard]] auto device_state() const -> kq_state*;
    void mark_device_modified() const;

private:
    std::size_t n_qubits_;
    std::size_t n_states_;
    Precision precision_;
    mutable std::vector<std::complex<double>> host_;
    mutable std::shared_ptr<kq_state> device_;
    mutable bool host_valid_ {false};
    mutable bool device_valid_ {false};
    mutable std::optional<std::size_t> basis_;   // the state is exactly |basis_>, not yet materialised

    void ensure_host_() const;
    void check_index_(std::size_t index) const;
};

auto almost_eq(const Statevector& left, const Statevector& right,
               double tolerance_sq = COMPLEX_ALMOST_EQ_TOLERANCE_SQ) noexcept -> bool;
auto inner_product(const Statevector& bra_state, const Statevector& ket_state) -> std::complex<double>;
auto inner_product_norm_squared(const Statevector& left, const Statevector& right) -> double;
auto tensor_product(const Statevector& left, const Statevector& right) -> Statevector;

}  // namespace ket
