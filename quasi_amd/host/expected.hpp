// expected.hpp -- a small value-or-error type for the host code.
//
// The reference host returns std::expected<T, E> (C++23, e.g.
// src/quasi/plugin/dynamic_library.hpp:56); the GCC 11 standard library of
// this toolchain has no <expected>, so the host uses this stand-in with the
// same observers (has_value, operator bool, operator*, operator->, value,
// error) and Q::unexpected(e) to build the error case.
#pragma once

#include <type_traits>
#include <utility>
#include <variant>

namespace Q {

template <class E>
struct unexpected_value {
    E error;
};

template <class E>
unexpected_value<std::decay_t<E>> unexpected(E&& e) {
    return {std::forward<E>(e)};
}

template <class T, class E>
class expected {
public:
    expected(T value) : v_(std::in_place_index<0>, std::move(value)) {}
    template <class G>
    expected(unexpected_value<G> u) : v_(std::in_place_index<1>, E(std::move(u.error))) {}

    bool has_value() const noexcept { return v_.index() == 0; }
    explicit operator bool() const noexcept { return has_value(); }
    T& operator*() & { return std::get<0>(v_); }
    const T& operator*() const& { return std::get<0>(v_); }
    T&& operator*() && { return std::get<0>(std::move(v_)); }
    T* operator->() { return &std::get<0>(v_); }
    const T* operator->() const { return &std::get<0>(v_); }
    T& value() & { return std::get<0>(v_); }
    const E& error() const { return std::get<1>(v_); }

private:
    std::variant<T, E> v_;
};

// expected<void, E>: success or an error.
template <class E>
class expected<void, E> {
public:
    expected() = default;
    template <class G>
    expected(unexpected_value<G> u) : err_(E(std::move(u.error))), ok_(false) {}

    bool has_value() const noexcept { return ok_; }
    explicit operator bool() const noexcept { return ok_; }
    const E& error() const { return err_; }

private:
    E err_{};
    bool ok_ = true;
};

}  // namespace Q
