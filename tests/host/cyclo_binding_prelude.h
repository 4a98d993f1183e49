// The few declarations of the reference program that the transform binding
// uses, given just enough body to compile and run the binding here: VSIZE
// (cyclo/src/engine.h:19), the integer typedefs and vint32
// (cyclo/src/transform.h:21), the engine the transform receives, pio::error
// (cyclo/src/pio.h:111) and ilog2_32.  The reference's own headers need
// OpenCL and are not used.
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#define VSIZE 64

typedef int32_t int32;
typedef uint32_t uint32;
typedef uint64_t uint64;
typedef std::array<uint32, VSIZE> vint32;

class engine {};

struct pio
{
	static void error(const std::string &str, const bool fatal = false)
	{
		fprintf(stderr, "error: %s\n", str.c_str());
		if (fatal) exit(1);
	}
};

inline int ilog2_32(const uint32_t n) { return 31 - __builtin_clz(n); }
