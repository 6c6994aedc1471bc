"""Oracle: CPU restatement of cunuSHT's nuSHT algorithm, FP64.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may import it.  The
product package (cunusht_amd) never imports, calls or links anything here.

Contents
  sht.py       numpy spin-weighted Y_lm (Wigner-d recursion), direct summation
               and its exact adjoint (small lmax).
  almutil.py   alm layout, seeded synalm, reference dot-product (myalmdot).
  geometry.py  GL / CC rings and the reference's good_size rule.
  src/*.c      C/OpenMP FP64 restatement of the reference CPU backend's stages
               (ring Legendre transforms, ES nuFFT, pointing), built into
               oracle/liboracle.so by oracle/Makefile.
  pipeline.py  the CPU_DUCCnufft_transformer algorithm restated on top of those.

Parity status of each function is listed in the header of the module that
defines it.
"""
