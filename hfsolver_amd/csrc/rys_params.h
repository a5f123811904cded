// Generated by gen_rys_tables.py - do not edit.
// Piecewise-Chebyshev Rys nodes/weights on T in [0, 60), Gauss-Hermite beyond.
#pragma once
#define RYS_NMAX 5
#define RYS_NINT 60
#define RYS_DEG 13
#define RYS_H 1
#define RYS_TMAX 60
#define RYS_CHEB_SIZE 25200
