! fvwb_abi_check — checks the Fortran view of the C ABI without touching a GPU:
! the size of the bind(C) parameter struct, the reference set-up values filled
! by fvwb_params_reference_2d, and that fvwb_create rejects an invalid set-up
! with FVWB_ERR_ARG and a message; likewise for the 1D parameters, the 2D DG
! structs, the reference DG set-up and its host Gauss rule.  Prints one line
! per field.
program fvwb_abi_check
  use iso_c_binding
  use fvwb_iface
  implicit none
  type(fvwb_params) :: p
  type(fvwb1d_params) :: p1
  type(fvwb_dg_params) :: pd
  type(fvwb_dgq_params) :: pq
  real(c_double) :: xq(3), wq(3)
  type(c_ptr) :: h
  integer(c_int) :: rc
  write(*, '(a,i0)') 'sizeof ', c_sizeof(p)
  rc = fvwb_params_reference_2d(p, 300_c_int32_t, 200_c_int32_t, 2_c_int32_t)
  write(*, '(a,i0)') 'status ', rc
  write(*, '(a,8(1x,i0))') 'ints', p%nvar, p%nx, p%ny, p%bc, p%recon, p%flux, p%integrator, p%precision
  write(*, '(a,7(1x,z16.16))') 'doubles', p%gamma, p%boxlen_x, p%boxlen_y, p%cfl, p%rho0, p%p0, p%g
  p%nx = 2
  rc = fvwb_create(h, p)
  write(*, '(a,i0)') 'create_bad ', rc
  write(*, '(a,a)') 'message ', fvwb_message(fvwb_last_error(c_null_ptr))
  write(*, '(a,i0)') 'sizeof1d ', c_sizeof(p1)
  rc = fvwb1d_params_reference(p1, 128_c_int32_t, 2_c_int32_t)
  write(*, '(a,6(1x,i0))') 'ints1d', p1%nvar, p1%nx, p1%bc, p1%recon, p1%integrator, p1%nequilibrium
  write(*, '(a,4(1x,z16.16))') 'doubles1d', p1%gamma, p1%boxlen, p1%cfl, p1%g
  write(*, '(a,i0)') 'sizeofdg ', c_sizeof(pd)
  write(*, '(a,i0)') 'sizeofdgq ', c_sizeof(pq)
  rc = fvwb_dgq_params_reference(pq)
  write(*, '(a,10(1x,i0))') 'intsdgq', pq%nx, pq%ny, pq%order, pq%flux_type, pq%source, pq%grad_phi_case, &
       pq%bc, pq%ninit, pq%solver, pq%use_limiter
  write(*, '(a,5(1x,z16.16))') 'doublesdgq', pq%gamma, pq%boxlen_x, pq%boxlen_y, pq%cfl, pq%eps
  pq%ny = 9
  rc = fvwb_dgq_create(h, pq)
  write(*, '(a,i0)') 'createdgq_bad ', rc
  rc = fvwb_dgq_gl_quadrature(3_c_int32_t, xq, wq)
  write(*, '(a,6(1x,z16.16))') 'glq3', xq, wq
end program fvwb_abi_check
