! spk_fortran_check.f90 -- a Fortran host that drives the GPU solve loop
! through the reference-style module API (lib_sie_tri_mlfmm_gpu), the way
! run_ml_fmm_GMRES_tri (src/lib_sie_solver_gmres_interface.f90:38-44) does.
!
! usage: spk_fortran_check pp_file tt_file out_file
! Reads the pp_/tt_ text files (one node / one triangle per line), runs the
! precalculation, one matvec of x = (1 + i*m/N) and a GMRES solve, and writes
! N, the matvec result and the solution (re im per line) to out_file.
program spk_fortran_check
   use lib_sie_tri_mlfmm_gpu
   implicit none
   character(len=512) :: fpp, ftt, fout
   double complex, allocatable :: x(:), y(:), b(:), xs(:)
   integer :: i, ios, its
   real(8) :: be

   call get_command_argument(1, fpp)
   call get_command_argument(2, ftt)
   call get_command_argument(3, fout)

   ! PMCHWT, calc_p(7) = 2, ng = 3, s_opt = 250, K = 19; lambda 633 nm, n = 1.5
   call lib_sie_ml_fmm_gpu_init(633d-9, (1d0, 0d0), (2.25d0, 0d0), 1, 2, 3, 250, 19)
   ! the reference's argument-free forms: the mesh files and the plane wave
   ! are registered first, as its data container and illumination_p hold them
   call lib_sie_ml_fmm_set_mesh_files(trim(fpp), trim(ftt))
   call lib_sie_ml_fmm_set_illumination([0d0, 0d0, -1d0], [1d0, 0d0, 0d0])
   call lib_sie_ml_fmm_precalculation_tri()

   allocate(x(2*m_pairs))
   do i = 1, 2*m_pairs
      x(i) = dcmplx(1d0, dble(i) / dble(2*m_pairs))
   end do
   call lib_sie_ml_fmm_calculate_vector_b_tri(x, y)

   call lib_sie_ml_fmm_get_vector_b_tri(b)
   call lib_sie_ml_fmm_get_init_vector_x_tri(xs, 0d0)
   call lib_sie_ml_fmm_gmres_run(b, xs, 300, 0, 1d-6, 1, its, be)

   open(unit=83, file=fout, action='write', status='replace', iostat=ios)
   write(83, *) m_pairs, its, be
   do i = 1, 2*m_pairs
      write(83, '(2es26.17)') real(y(i), 8), aimag(y(i))
   end do
   do i = 1, 2*m_pairs
      write(83, '(2es26.17)') real(xs(i), 8), aimag(xs(i))
   end do
   close(83)
   call lib_sie_ml_fmm_gpu_finalize()

end program spk_fortran_check
