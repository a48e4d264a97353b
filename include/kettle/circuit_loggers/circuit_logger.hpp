#pragma once

#include <memory>
#include <stdexcept>
#include <variant>

#include "kettle/circuit/classical_register.hpp"
#include "kettle/state/statevector.hpp"

// Mid-circuit snapshots (circuit_loggers/*.hpp in the reference).
namespace ket
{

class StatevectorCircuitLogger
{
public:
    void add_statevector(Statevector statevector) { statevector_ = std::make_shared<Statevector>(std::move(statevector)); }
    [[nodiscard]] auto statevector() const -> const Statevector&
    {
        if (!statevector_) {
            throw std::runtime_error {"ERROR: Cannot access statevector; it is not in the logger\n"};
        }
        return *statevector_;
    }

private:
    std::shared_ptr<Statevector> statevector_;
};

class ClassicalRegisterCircuitLogger
{
public:
    void add_classical_register(ClassicalRegister cregister)
    {
        cregister_ = std::make_shared<ClassicalRegister>(std::move(cregister));
    }
    [[nodiscard]] auto classical_register() const -> const ClassicalRegister&
    {
        if (!cregister_) {
            throw std::runtime_error {"ERROR: Cannot access classical register; it is not in the logger\n"};
        }
        return *cregister_;
    }

private:
    std::shared_ptr<ClassicalRegister> cregister_;
};

class CircuitLogger
{
public:
    CircuitLogger(StatevectorCircuitLogger logger) : logger_ {std::move(logger)} {}             // NOLINT
    CircuitLogger(ClassicalRegisterCircuitLogger logger) : logger_ {std::move(logger)} {}       // NOLINT

    [[nodiscard]] auto is_statevector_circuit_logger() const -> bool
    {
        return std::holds_alternative<StatevectorCircuitLogger>(logger_);
    }
    [[nodiscard]] auto is_classical_register_circuit_logger() const -> bool
    {
        return std::holds_alternative<ClassicalRegisterCircuitLogger>(logger_);
    }
    [[nodiscard]] auto get_statevector_circuit_logger() const -> const StatevectorCircuitLogger&
    {
        return std::get<StatevectorCircuitLogger>(logger_);
    }
    [[nodiscard]] auto get_classical_register_circuit_logger() const -> const ClassicalRegisterCircuitLogger&
    {
        return std::get<ClassicalRegisterCircuitLogger>(logger_);
    }

private:
    std::variant<StatevectorCircuitLogger, ClassicalRegisterCircuitLogger> logger_;
};

}  // namespace ket
