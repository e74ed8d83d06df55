// Morpheus_AMD_Spaces.hpp -- execution / memory spaces, layouts, the
// exception taxonomy and initialize/finalize of the MI355X drop-in for the
// Morpheus SpMV path.  Host C++17 over the C-ABI of include/morpheus_amd.h.
//
// Mirrors core/src/Morpheus_CustomBackend.hpp:96-181 (Serial, OpenMP, HIP,
// HostSpace, HIPSpace, DefaultExecutionSpace), core/src/
// Morpheus_GenericBackend.hpp:128 (Generic::HIP), core/src/
// Morpheus_Exceptions.hpp:32-74 and core/src/Morpheus_Core.hpp:74-91.
// Kokkos is not used: the space and layout tags are plain types, and a small
// Kokkos:: tag namespace keeps user code that names Kokkos::HostSpace or
// Kokkos::LayoutRight compiling (define MORPHEUS_AMD_NO_KOKKOS_SHIM to drop it).
#pragma once

#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>
#include <string>

#include "morpheus_amd.h"

namespace Morpheus {

// ------------------------------------------------------------ exceptions
class Exception : public std::exception {
 public:
  Exception(const Exception& e) : message(e.message) {}
  Exception(const std::string& m) : message(m) {}
  ~Exception() throw() {}
  const char* what() const throw() { return message.c_str(); }

 protected:
  std::string message;
};

class NotImplementedException : public Exception {
 public:
  template <typename M>
  NotImplementedException(const M& fn)
      : Exception(std::string("NotImplemented: ") + fn + std::string(" not yet implemented.")) {}
};

class IOException : public Exception {
 public:
  template <typename M>
  IOException(const M& msg) : Exception(std::string("IO: ") + msg) {}
};

class RuntimeException : public Exception {
 public:
  template <typename M>
  RuntimeException(const M& msg) : Exception(msg) {}
};

class InvalidInputException : public Exception {
 public:
  template <typename M>
  InvalidInputException(const M& msg) : Exception(msg) {}
};

class FormatConversionException : public Exception {
 public:
  template <typename M>
  FormatConversionException(const M& msg) : Exception(msg) {}
};

namespace Impl {
// status of a C-ABI call -> the matching Morpheus exception
inline void check(int status, const char* what) {
  if (status == MPH_OK) return;
  const std::string msg = std::string(what) + ": " + mph_last_error();
  switch (status) {
    case MPH_ERR_NOT_IMPLEMENTED: throw NotImplementedException(msg);
    case MPH_ERR_INVALID_INPUT: throw InvalidInputException(msg);
    case MPH_ERR_FORMAT_CONVERSION: throw FormatConversionException(msg);
    default: throw RuntimeException(msg);
  }
}
}  // namespace Impl

// ------------------------------------------------------------ formats
// core/src/Morpheus_FormatsRegistry.hpp:56-64
enum formats_e {
  COO_FORMAT = 0,
  CSR_FORMAT,
  DIA_FORMAT,
  ELL_FORMAT,
  HYB_FORMAT,
  HDC_FORMAT,
  NFORMATS
};

// ------------------------------------------------------------ layouts
struct LayoutLeft {};   // column-major (the device layout)
struct LayoutRight {};  // row-major (the host default)

// ------------------------------------------------------------ spaces
struct Serial;
struct OpenMP;
struct HIP;

struct HostSpace {
  static constexpr bool is_device = false;
  using execution_space = Serial;
  using default_layout = LayoutRight;
};

struct HIPSpace {
  static constexpr bool is_device = true;
  using execution_space = HIP;
  using default_layout = LayoutLeft;
};

// Execution spaces of the custom backend.  Serial / OpenMP algorithms on
// host containers are served by the same GPU kernels through a device round
// trip; the HIP space runs them on resident data.
struct Serial {
  using memory_space = HostSpace;
  static constexpr bool is_device = false;
};
struct OpenMP {
  using memory_space = HostSpace;
  static constexpr bool is_device = false;
};
struct HIP {
  using memory_space = HIPSpace;
  static constexpr bool is_device = true;
};

namespace Custom {
using Serial = Morpheus::Serial;
using OpenMP = Morpheus::OpenMP;
using HIP = Morpheus::HIP;
using HostSpace = Morpheus::HostSpace;
using HIPSpace = Morpheus::HIPSpace;
using DefaultExecutionSpace = Morpheus::HIP;
using DefaultHostExecutionSpace = Morpheus::Serial;
}  // namespace Custom

// Morpheus::Generic::HIP dispatches to the same MI355X kernels.
namespace Generic {
struct HIP : Morpheus::HIP {};
struct Serial : Morpheus::Serial {};
struct OpenMP : Morpheus::OpenMP {};
using HostSpace = Morpheus::HostSpace;
using HIPSpace = Morpheus::HIPSpace;
}  // namespace Generic

using DefaultExecutionSpace = HIP;
using DefaultHostExecutionSpace = Serial;

template <class T>
struct is_device_space {
  static constexpr bool value = T::is_device;
};

// memory space named by a container property: a memory space, or an
// execution space standing for its memory space
template <class S, class = void>
struct memory_space_of {
  using type = S;
};
template <class S>
struct memory_space_of<S, std::void_t<typename S::memory_space>> {
  using type = typename S::memory_space;
};

// ------------------------------------------------------------ init
struct InitArguments {
  int dynamic_format = 0;
};

namespace Impl {
// core/src/impl/Morpheus_Core.cpp:92-140: --morpheus-format=INT / --format=INT
inline void parse_command_line_arguments(int& argc, char* argv[], InitArguments& args) {
  int i = 1;
  while (i < argc) {
    const char* a = argv[i];
    const char* v = nullptr;
    if (std::strncmp(a, "--morpheus-format=", 18) == 0) v = a + 18;
    else if (std::strncmp(a, "--format=", 9) == 0) v = a + 9;
    if (v) {
      char* end = nullptr;
      const long f = std::strtol(v, &end, 10);
      if (end == v || *end != '\0' || f < 0 || f >= NFORMATS)
        throw RuntimeException(std::string("Invalid format: ") + v);
      args.dynamic_format = (int)f;
      for (int j = i; j < argc - 1; j++) argv[j] = argv[j + 1];
      argc--;
      continue;
    }
    if (std::strcmp(a, "--morpheus-help") == 0) {
      std::cout << "--morpheus-format=INT  format of DynamicMatrix containers (0..5)\n";
    }
    i++;
  }
}
}  // namespace Impl

// core/src/Morpheus_Core.hpp:74-91.  The parsed arguments are returned
// through the InitArguments overload.
inline void initialize(int& argc, char* argv[], InitArguments& args) {
  Impl::parse_command_line_arguments(argc, argv, args);
  Impl::check(mph_initialize(-1), "Morpheus::initialize");
}
inline void initialize(int& argc, char* argv[]) {
  InitArguments args;
  initialize(argc, argv, args);
}
inline void initialize(InitArguments = InitArguments()) {
  Impl::check(mph_initialize(-1), "Morpheus::initialize");
}
inline void finalize() { mph_finalize(); }

}  // namespace Morpheus

#ifndef MORPHEUS_AMD_NO_KOKKOS_SHIM
namespace Kokkos {
using HostSpace = Morpheus::HostSpace;
using LayoutLeft = Morpheus::LayoutLeft;
using LayoutRight = Morpheus::LayoutRight;
using Serial = Morpheus::Serial;
using OpenMP = Morpheus::OpenMP;
using DefaultExecutionSpace = Morpheus::HIP;
using DefaultHostExecutionSpace = Morpheus::Serial;
namespace Experimental {
using HIP = Morpheus::HIP;
using HIPSpace = Morpheus::HIPSpace;
}  // namespace Experimental
}  // namespace Kokkos
#endif
