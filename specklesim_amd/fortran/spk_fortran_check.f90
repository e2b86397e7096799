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
   real(8), allocatable :: pp(:, :)
   integer, allocatable :: tt(:, :)
   double complex, allocatable :: x(:), y(:), b(:), xs(:)
   integer :: np, nt, i, ios, its
   real(8) :: buf(3), be
   integer :: ibuf(3)

   call get_command_argument(1, fpp)
   call get_command_argument(2, ftt)
   call get_command_argument(3, fout)

   np = count_lines(fpp)
   nt = count_lines(ftt)
   allocate(pp(3, np), tt(3, nt))
   open(unit=81, file=fpp, status='old', action='read')
   do i = 1, np
      read(81, *) buf
      pp(:, i) = buf
   end do
   close(81)
   open(unit=82, file=ftt, status='old', action='read')
   do i = 1, nt
      read(82, *) ibuf
      tt(:, i) = ibuf
   end do
   close(82)

   ! PMCHWT, calc_p(7) = 2, ng = 3, s_opt = 250, K = 19; lambda 633 nm, n = 1.5
   call lib_sie_ml_fmm_gpu_init(633d-9, (1d0, 0d0), (2.25d0, 0d0), 1, 2, 3, 250, 19)
   call lib_sie_ml_fmm_precalculation_tri(pp, tt)

   allocate(x(2*m_pairs))
   do i = 1, 2*m_pairs
      x(i) = dcmplx(1d0, dble(i) / dble(2*m_pairs))
   end do
   call lib_sie_ml_fmm_calculate_vector_b_tri(x, y)

   call lib_sie_ml_fmm_get_vector_b_tri(b, [0d0, 0d0, -1d0], [1d0, 0d0, 0d0])
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

contains

   integer function count_lines(fname)
      character(len=*), intent(in) :: fname
      integer :: st
      character(len=1024) :: line
      count_lines = 0
      open(unit=80, file=fname, status='old', action='read')
      do
         read(80, '(a)', iostat=st) line
         if (st /= 0) exit
         if (len_trim(line) > 0) count_lines = count_lines + 1
      end do
      close(80)
   end function
end program spk_fortran_check
