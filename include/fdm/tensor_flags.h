// Axis flags of fdm's solver templates (the interface of reference
// src/tensor.h:12-62: tensor_flag, tensor_flags<...>, has_tensor_flag).
//
// When the drop-in headers are used inside the reference tree, define
// FDMGPU_WITH_FDM_TENSOR so the reference's own tensor.h provides these types
// (both cannot be defined in one program).
#pragma once

#ifdef FDMGPU_WITH_FDM_TENSOR
#include "tensor.h"
#else

namespace fdm {

enum struct tensor_flag : int { none = 0, periodic = 1 };

template <tensor_flag... Args>
struct tensor_flags;

template <tensor_flag Head, tensor_flag... Tail>
struct tensor_flags<Head, Tail...> {
    using tail = tensor_flags<Tail...>;
    static constexpr tensor_flag head = Head;
};

template <>
struct tensor_flags<> {
    using tail = tensor_flags<>;
    static constexpr tensor_flag head = tensor_flag::none;
};

constexpr bool has_tensor_flag(tensor_flag flags, tensor_flag flag) {
    return (static_cast<int>(flags) & static_cast<int>(flag)) != 0;
}

}  // namespace fdm

#endif
