// Minimal assertion helpers for the C++ mirror tests (exit status 1 on failure).
#pragma once
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

inline void check_fail(const std::string& m) {
  std::cerr << "CHECK FAILURE: " << m << "\n";
  std::exit(1);
}
#define CHECK_TRUE(c) \
  do { if (!(c)) check_fail("expected " #c); } while (false)
#define CHECK_NEAR(a, b, tol)                                                              \
  do {                                                                                     \
    const double a_ = static_cast<double>(a), b_ = static_cast<double>(b);                 \
    if (!(std::fabs(a_ - b_) <= static_cast<double>(tol)))                                 \
      check_fail(std::string(#a " vs " #b ": ") + std::to_string(a_) + " / " +             \
                 std::to_string(b_));                                                      \
  } while (false)
