#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <string>
#include <unordered_map>

// Parameters of parameterised gates (include/kettle/parameter/parameter.hpp in the reference).
namespace ket::param
{

struct parameterized
{
    explicit parameterized() = default;
};

constexpr inline auto PARAMETER_ID_SIZE = std::size_t {16};
using ParameterID = std::array<std::uint8_t, PARAMETER_ID_SIZE>;

class Parameter
{
public:
    Parameter(std::string name, const ParameterID& id);
    Parameter(std::string name, int seed);
    explicit Parameter(std::string name);

    [[nodiscard]] auto name() const -> const std::string& { return name_; }
    [[nodiscard]] auto id() const -> const ParameterID& { return id_; }

private:
    std::string name_;
    ParameterID id_;
};

struct ParameterIdHash
{
    auto operator()(const ParameterID& id) const noexcept -> std::size_t;
};

struct ParameterData
{
    std::optional<double> value;
    std::string name;
    std::size_t count;
};

using ParameterDataMap = std::unordered_map<ParameterID, ParameterData, ParameterIdHash>;

}  // namespace ket::param
