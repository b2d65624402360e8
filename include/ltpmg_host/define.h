// define.h -- common includes and scalar typedefs of the host API
// (same names as the reference's include/define.h:39-55).  gflags/glog come
// from ltpmg_flags.h; the JSON schema loader is not part of this build, so
// there is no nlohmann include or `json` alias.
#pragma once

#ifndef LTPMG_DEFINE
#define LTPMG_DEFINE

#include <sys/types.h>

#include <algorithm>
#include <any>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <future>
#include <iostream>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <tuple>
#include <typeinfo>
#include <unordered_set>
#include <variant>
#include <vector>

#include "ltpmg_flags.h"

typedef int8_t INT8;
typedef int16_t INT16;
typedef int32_t INT32;
typedef int64_t INT64;

typedef uint8_t UINT8;
typedef uint16_t UINT16;
typedef uint32_t UINT32;
typedef uint64_t UINT64;

typedef char CHAR;
typedef u_char UCHAR;

typedef float FLOAT;
typedef double DOUBLE;

typedef std::string STRING;

typedef bool BOOL;

#endif
