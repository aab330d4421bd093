"""Closed-form values that pin the oracle and the device path independently
of either (shared by the CPU and GPU tests)."""


def bmskel_node1():
    """Node 1 of the reference's tests/data/bmskel deck after its one
    Crank-Nicolson step.  Line element of length 1: M = rho*c/6 [[2,1],[1,2]],
    K = lambda [[1,-1],[-1,1]] with rho*c = 2000*900 (a.msp), lambda = 5,
    dt = 1, theta = 0.5, T_n = (273, 0) (the initial condition reaches POINT0
    only).  Row 1 keeps its column 0 under row-only elimination
    (CSparseMatrix::Diagonize, FEM/matrix_class.cpp:1663), so
    T_1 = (RHS_1 - LHS_10 * 274) / LHS_11 with LHS = M/dt + K/2 and
    RHS = (M/dt - K/2) T_n (AssembleMixedHyperbolicParabolicEquation,
    FEM/fem_ele_std.cpp:8504)."""
    m, k = 2000.0 * 900.0 / 6.0, 5.0
    lhs10, lhs11 = m - 0.5 * k, 2 * m + 0.5 * k
    rhs1 = (m + 0.5 * k) * 273.0
    return (rhs1 - lhs10 * 274.0) / lhs11
