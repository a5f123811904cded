// Generated by gen_rys_tables.py - do not edit.
// Piecewise-polynomial Rys nodes/weights on T in [0, 60), Gauss-Hermite beyond.
#pragma once
#define RYS_NMAX 5
#define RYS_NINT 480
#define RYS_DEG 7
#define RYS_H 0.125
#define RYS_TMAX 60
#define RYS_POLY_SIZE 115200
#define RYS_DEGF 6
#define RYS_CHEBF_SIZE 100800
#define BOYS_DT 0.125
#define BOYS_INV_DT 8
#define BOYS_TMAX 46
#define BOYS_K 9
#define BOYS_NCOL 12
#define BOYS_NPT 369
#define BOYS_SIZE 4428
#define BOYSF_K 5
#define BOYSF_NCOL 8
#define BOYSF_SIZE 2952
