"""CPU oracle for SuperNormal's training hot path.

TEST INFRASTRUCTURE ONLY: the checker that the HIP kernels are compared
against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this package; the product (supernormal_amd) never does.

Parts:
  sn_oracle.c / c_oracle.py   bit-exact sequential restatements (marching,
                              pack_info, visibility, patch weights, occupancy,
                              hash-grid corner selection), built by the Makefile
                              into oracle/_build/liboracle_sn.so
  hashgrid_oracle.py          hash-grid encode in fp64 torch (autograd gives
                              backward and double backward)
  field_oracle.py             SDF network (weight-normed Softplus MLP) in torch
  train_oracle.py             one training step of the reference restated in
                              torch on the CPU (the "plain method on CPU")
"""
