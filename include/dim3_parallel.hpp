// dim3_parallel.hpp — drop-in counterpart of the reference's
// dim3_parallel.hpp (dim3_parallel.hpp:1-10): callers of
// DIM3_parallel<...>::doJoinProject_parallel include this file.
//
// DIM3_parallel itself is declared in dim3.hpp (the query runs on the GPU,
// so there is no OpenMP team); THREAD_NUM keeps the reference's value for
// callers that size their own work by it.
#pragma once

#ifndef THREAD_NUM
#define THREAD_NUM 8
#endif

#include "dim3.hpp"
