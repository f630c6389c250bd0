"""Oracle for the DelFEM FEM hot path: TEST INFRASTRUCTURE ONLY.

- `cpu_ref.c` (built to `oracle/libcpu_ref.so`): plain-C restatement of the reference's arithmetic.
- `_ref/libdfm_ref.so` (built by `make -C oracle ref` from the reference's own sources, in the container that
  holds them): the reference itself behind a small driver (`ref_capi.cpp`), used to pin the restatement.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this package. The product
(`delfem_amd`) never does.
"""
