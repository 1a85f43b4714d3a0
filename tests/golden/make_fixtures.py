"""Write tests/golden/reference_known_answers.json.

The inputs are the hand-written known-answer cases of the reference's own
tests (file:line given per case); the expected outputs are computed here by
closed form with Python integers (exact product then mod, identity, the
stack semantics of permutations.jl), not by any code under test.

    python tests/golden/make_fixtures.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def matmul(A, B):
    return [[sum(A[i][t] * B[t][j] for t in range(len(B))) for j in range(len(B[0]))] for i in range(len(A))]


def mod(X, p):
    return [[x % p for x in row] for row in X]


def swap_rows(X, stack):
    X = [list(r) for r in X]
    for a, b in stack:
        X[a - 1], X[b - 1] = X[b - 1], X[a - 1]
    return X


def swap_cols(X, stack):
    X = [list(r) for r in X]
    for a, b in stack:
        for r in X:
            r[a - 1], r[b - 1] = r[b - 1], r[a - 1]
    return X


def main():
    cases = []
    A = [[1, 2, 3], [4, 5, 6]]
    B = [[7, 8], [9, 10], [0, 1]]
    cases.append(dict(name="matmul_2x3x2_mod11", src="test/CuModMatrix/matmul_operations_test.jl:14-41",
                      A=A, B=B, p=11, C=mod(matmul(A, B), 11)))
    cases.append(dict(name="matmul_2x3x2_override_mod7", src="test/CuModMatrix/matmul_operations_test.jl:54-67",
                      A=A, B=B, p=7, C=mod(matmul(A, B), 7)))
    B2 = [[7, 8], [9, 10], [11, 12]]
    for p in (9, 3):
        cases.append(dict(name=f"matmul_inplace_mod{p}", src="test/CuModMatrix/matmul_operations_test.jl:79-126",
                          A=A, B=B2, p=p, C=mod(matmul(A, B2), p)))
    A3 = [[1, 2, 3], [4, 5, 6], [7, 8, 9]]
    B3 = [[9, 8, 7], [6, 5, 4], [3, 2, 1]]
    cases.append(dict(name="matmul_3x3_mod11", src="test/CuModMatrix/basic_operations_test.jl:18-71; "
                      "test/CuModMatrix/inplace_operations_test.jl:99-107", A=A3, B=B3, p=11,
                      C=mod(matmul(A3, B3), 11)))
    G = [[1, 2, 3], [0, 1, 4], [5, 6, 0]]
    cases.append(dict(name="power_G2_mod11", src="test/CuModMatrix/basic_operations_test.jl:161-194",
                      A=G, B=G, p=11, C=mod(matmul(G, G), 11)))

    perm = dict(name="permutation_stack_2_3", src="test/CuModMatrix/permutation_test.jl:3-55",
                A=A3, p=11, stack=[[2, 3]],
                col_perm=swap_cols(A3, [(2, 3)]), row_perm=swap_rows(A3, [(2, 3)]))

    derham = [
        [0, 0, 0, 0, 0, 0, 0, -3, 0, 0],
        [0, 0, 0, 0, 0, 0, 0, 0, -3, 0],
        [0, 1, 0, 0, 0, 0, 0, 0, 0, 0],
        [0, 0, 1, 0, 0, 0, 0, 0, 0, 0],
        [0, -2, 0, 0, 0, 0, 0, 0, 0, -3],
        [0, 0, -2, 0, 2, 0, 0, 0, 0, 0],
        [0, 0, 0, 1, 0, 2, 0, 0, 0, 0],
        [0, -3, 0, -2, 0, 0, 0, -1, 0, 0],
        [0, 0, -3, 0, 0, 0, 2, 0, -1, 0],
        [1, 0, 0, -3, 0, 0, 0, 0, 0, -1],
    ]
    elim = [
        dict(name="de_rham_10x10_mod7", src="test/CuModMatrix/de_rham_test.jl:3-37",
             A=derham, p=7, rank=10,
             # pivot stack of the src restatement (largest residue, lowest
             # row); a cross-check value, not a closed form
             row_swaps=[[1, 10], [2, 5], [3, 6], [4, 8], [5, 9], [6, 7], [7, 8], [8, 10]], col_swaps=[]),
        dict(name="identity_6x6_mod11", src="test/CuModMatrix/subst_test.jl:8,33-48",
             A=[[1 if i == j else 0 for j in range(6)] for i in range(6)], p=11, rank=6,
             row_swaps=[], col_swaps=[]),
    ]
    # stripe_mul_test.jl:5-50 injects residues p-5 and p-23 into random data
    stripe = [dict(name="stripe_mul_inject", src="test/CuModMatrix/stripe_mul_test.jl:16-18,35-37",
                   p=11 ** 3, size=100, inject=[[5, 5, "A", 11 ** 3 - 5], [95, 95, "B", 11 ** 3 - 23]]),
              dict(name="stripe_mul_mod2^11", src="test/CuModMatrix/stripe_mul_test.jl:2-13", p=2 ** 11,
                   size=100, inject=[])]
    out = dict(matmul=cases, permutation=perm, elimination=elim, stripe=stripe)
    with open(os.path.join(HERE, "reference_known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
