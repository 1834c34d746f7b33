"""psfun_amd -- MI355X (gfx950) implementation of PSFUN's f(A)x hot path.

Host-side mirror of PSFUN's operator interface over the C-ABI of
libpsfun_hip.so (include/psfun_hip.h):

  psb            -- single-process PSBLAS surface (CSR matrix, vectors, primitives)
  serial         -- psfun_d_serial (the f(H) step)
  krylov         -- psfun_d_krylov (LANCZOS / ARNOLDI apply), psb_krylov (shifted CG)
  quadrature     -- psfun_z_quadrature, hhtmethod1, Talbot rule, quadrature apply
  batched        -- many independent f(A)x in one persistent launch (+ autotuner)
  autograd       -- torch.autograd.Function for exp(-tA)X with adjoint kernels
  matgen         -- seeded synthetic inputs and closed-form references
"""
__all__ = ["psb", "serial", "krylov", "matgen"]
