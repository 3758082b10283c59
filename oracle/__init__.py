"""CPU oracle for unitair's gate-application path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

- unitair_np    numpy restatement of the reference algorithm (the checker)
- closed_form   ctypes wrapper of closed_form.c (index-definition restatement)
- unitair_torch torch CPU restatement of the reference's op sequence (timed
                CPU baseline; autograd gradients for the adjoint checks)
"""
