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
#define BOYS_DT 0.03125
#define BOYS_INV_DT 32
#define BOYS_TMAX 46
#define BOYS_K 7
#define BOYS_NCOL 10
#define BOYS_NPT 1473
#define BOYS_SIZE 14730
#define BOYS1_KF 5
#define BOYS1_SIZE 5892
#define BOYSF_K 5
#define BOYSF_NCOL 8
#define BOYSF_SIZE 11784
