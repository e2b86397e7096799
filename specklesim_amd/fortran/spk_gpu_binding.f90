! spk_gpu_binding.f90 -- Fortran 2003 ISO_C_BINDING interface to
! libspeckle_gpu.so (include/speckle_gpu.h), and a drop-in module that keeps
! the API of the reference's lib_sie_tri_mlfmm_interface
! (src/lib_sie_tri_mlfmm_interface.f90:18-26) on top of it.
!
! Vectors are complex(8) arrays of length 2*m_pairs in the reference order
! [xe(1:m_pairs); xh(1:m_pairs)]; complex(8) is layout-compatible with the
! C side's interleaved (re, im) doubles.
module spk_gpu_binding
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: SPK_PMCHWT = 1, SPK_MCTF = 2, SPK_ICTF = 3, SPK_MCTF2 = 4, &
                                SPK_PMCHWT_S = 5, SPK_JMCFIE = 6, SPK_CNF = 7, SPK_MNMF = 8
   integer(c_int), parameter :: SPK_ORTH_MGS = 0, SPK_ORTH_IMGS = 1, SPK_ORTH_CGS = 2, &
                                SPK_ORTH_ICGS = 3

   type, bind(c) :: spk_params
      real(c_double) :: lambda
      real(c_double) :: eps_r1_re, eps_r1_im, eps_r2_re, eps_r2_im
      integer(c_int) :: formulation, normal_convention, quad_points, tree_s_opt
      integer(c_int) :: truncation_number, precision, device
   end type

   type, bind(c) :: spk_gmres_params
      integer(c_int) :: max_iterations, restart
      real(c_double) :: tolerance
      integer(c_int) :: orthogonalization, use_initial_guess, refinement_steps
      real(c_double) :: refinement_tolerance
      integer(c_int) :: preconditioning, inner_iterations
   end type

   type, bind(c) :: spk_gmres_info
      integer(c_int) :: info, iterations
      real(c_double) :: arnoldi_backward_error, backward_error
      integer(c_int) :: matvecs
      real(c_double) :: seconds
      real(c_double) :: true_backward_error
   end type

   interface
      integer(c_int) function spk_create(params, handle) bind(c, name="spk_create")
         import :: c_int, c_ptr, spk_params
         type(spk_params), intent(in) :: params
         type(c_ptr), intent(out) :: handle
      end function
      subroutine spk_destroy(handle) bind(c, name="spk_destroy")
         import :: c_ptr
         type(c_ptr), value :: handle
      end subroutine
      type(c_ptr) function spk_last_error() bind(c, name="spk_last_error")
         import :: c_ptr
      end function
      integer(c_int) function spk_setup_tri(handle, np, pp, nt, tt) bind(c, name="spk_setup_tri")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         integer(c_int), value :: np, nt
         real(c_double), intent(in) :: pp(*)
         integer(c_int), intent(in) :: tt(*)
      end function
      integer(c_int) function spk_num_edges(handle, n) bind(c, name="spk_num_edges")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int64_t), intent(out) :: n
      end function
      integer(c_int) function spk_matvec(handle, x, y) bind(c, name="spk_matvec")
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: handle
         complex(c_double_complex), intent(in) :: x(*)
         complex(c_double_complex), intent(out) :: y(*)
      end function
      integer(c_int) function spk_rhs_plane_wave(handle, k_in, e_in, b) &
            bind(c, name="spk_rhs_plane_wave")
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: handle
         real(c_double), intent(in) :: k_in(3), e_in(3)
         complex(c_double_complex), intent(out) :: b(*)
      end function
      ! 8-node quadrilaterals: the fill of lib_sie_quad_solver_normal
      integer(c_int) function spk_quad_create(params, handle) bind(c, name="spk_quad_create")
         import :: c_int, c_ptr, spk_params
         type(spk_params), intent(in) :: params
         type(c_ptr), intent(out) :: handle
      end function
      subroutine spk_quad_destroy(handle) bind(c, name="spk_quad_destroy")
         import :: c_ptr
         type(c_ptr), value :: handle
      end subroutine
      integer(c_int) function spk_quad_setup(handle, lx, ly, nx, ny, z) bind(c, name="spk_quad_setup")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), value :: lx, ly
         integer(c_int), value :: nx, ny
         real(c_double), intent(in) :: z(*)
      end function
      integer(c_int) function spk_quad_sizes(handle, n_edges, n_elements, n_nodes) &
            bind(c, name="spk_quad_sizes")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), intent(out) :: n_edges, n_elements, n_nodes
      end function
      integer(c_int) function spk_quad_edge_element(handle, n, s, t, out) &
            bind(c, name="spk_quad_edge_element")
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: handle
         integer(c_int), value :: n
         integer(c_int), intent(in) :: s(*), t(*)
         complex(c_double_complex), intent(out) :: out(*)
      end function
      integer(c_int) function spk_quad_fill(handle, mm) bind(c, name="spk_quad_fill")
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: handle
         complex(c_double_complex), intent(out) :: mm(*)
      end function
      ! get_vector_b_quad with a plane wave (src/lib_sie_quad_calculation_mod.f90:91-138)
      integer(c_int) function spk_quad_rhs_plane_wave(handle, k_in, e_in, b) &
            bind(c, name="spk_quad_rhs_plane_wave")
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: handle
         real(c_double), intent(in) :: k_in(3), e_in(3)
         complex(c_double_complex), intent(out) :: b(*)
      end function
      integer(c_int) function spk_gmres(handle, b, x, p, info) bind(c, name="spk_gmres")
         import :: c_int, c_ptr, c_double_complex, spk_gmres_params, spk_gmres_info
         type(c_ptr), value :: handle
         complex(c_double_complex), intent(in) :: b(*)
         complex(c_double_complex), intent(inout) :: x(*)
         type(spk_gmres_params), intent(in) :: p
         type(spk_gmres_info), intent(out) :: info
      end function
   end interface

contains

   ! stop with the library's message on a non-zero status (the reference
   ! prints and stops on its own errors, src/lib_sie_math.f90:669-674)
   subroutine spk_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      integer :: i
      character(len=512) :: text
      if (rc == 0) return
      call c_f_pointer(spk_last_error(), msg, [512])
      text = ' '
      do i = 1, 512
         if (msg(i) == c_null_char) exit
         text(i:i) = msg(i)
      end do
      print *, 'libspeckle_gpu error in ', where, ': ', trim(text)
      stop 1
   end subroutine
end module spk_gpu_binding


! Drop-in for lib_sie_tri_mlfmm_interface's public routines: the GMRES host
! and the illumination loop call these names unchanged.
module lib_sie_tri_mlfmm_gpu
   use iso_c_binding
   use spk_gpu_binding
   implicit none
   private

   public :: lib_sie_ml_fmm_gpu_init
   public :: lib_sie_ml_fmm_precalculation_tri
   public :: lib_sie_ml_fmm_calculate_vector_b_tri
   public :: lib_sie_ml_fmm_get_vector_b_tri
   public :: lib_sie_ml_fmm_get_init_vector_x_tri
   public :: lib_sie_ml_fmm_gmres_run
   public :: lib_sie_ml_fmm_gpu_finalize
   public :: lib_sie_ml_fmm_set_mesh_files
   public :: lib_sie_ml_fmm_set_mesh
   public :: lib_sie_ml_fmm_set_illumination
   public :: m_pairs

   type(c_ptr), save :: handle = c_null_ptr
   integer, save :: m_pairs = 0
   ! the state the reference's routines read instead of taking arguments:
   ! the mesh of the data container (data_import_tri reads it from the
   ! pp_/tt_ files) and the plane wave of illumination_p
   real(8), allocatable, save :: mesh_pp(:, :)
   integer, allocatable, save :: mesh_tt(:, :)
   real(8), save :: illum_k(3) = 0d0, illum_e(3) = 0d0
   logical, save :: illum_set = .false.

contains

   ! the module globals the reference sets in main.f90 / the data containers
   subroutine lib_sie_ml_fmm_gpu_init(lambda, eps_r1, eps_r2, formulation, calc_p7, ng, &
                                      tree_s_opt, truncation_number)
      real(8), intent(in) :: lambda
      complex(8), intent(in) :: eps_r1, eps_r2
      integer, intent(in) :: formulation, calc_p7, ng, tree_s_opt, truncation_number
      type(spk_params) :: p
      p%lambda = lambda
      p%eps_r1_re = real(eps_r1, 8); p%eps_r1_im = aimag(eps_r1)
      p%eps_r2_re = real(eps_r2, 8); p%eps_r2_im = aimag(eps_r2)
      p%formulation = formulation
      p%normal_convention = calc_p7
      p%quad_points = ng
      p%tree_s_opt = tree_s_opt
      p%truncation_number = truncation_number
      p%precision = 0
      p%device = 0
      call spk_check(spk_create(p, handle), 'spk_create')
   end subroutine

   ! data_import_tri / read_pp_tt_file_tri
   ! (src/lib_sie_tri_data_container.f90:226-261): the mesh files, one node
   ! (x y z) or one triangle (three 1-based node numbers) per line
   subroutine lib_sie_ml_fmm_set_mesh_files(pp_file, tt_file)
      character(len=*), intent(in) :: pp_file, tt_file
      integer :: np, nt, i
      real(8) :: buf(3)
      integer :: ibuf(3)
      np = count_lines(pp_file)
      nt = count_lines(tt_file)
      if (allocated(mesh_pp)) deallocate(mesh_pp)
      if (allocated(mesh_tt)) deallocate(mesh_tt)
      allocate(mesh_pp(3, np), mesh_tt(3, nt))
      open(unit=71, file=pp_file, status='old', action='read')
      do i = 1, np
         read(71, *) buf
         mesh_pp(:, i) = buf
      end do
      close(71)
      open(unit=72, file=tt_file, status='old', action='read')
      do i = 1, nt
         read(72, *) ibuf
         mesh_tt(:, i) = ibuf
      end do
      close(72)
   end subroutine

   ! the mesh given as arrays instead of files
   subroutine lib_sie_ml_fmm_set_mesh(pp, tt)
      real(8), intent(in) :: pp(:, :)
      integer, intent(in) :: tt(:, :)
      if (allocated(mesh_pp)) deallocate(mesh_pp)
      if (allocated(mesh_tt)) deallocate(mesh_tt)
      allocate(mesh_pp, source=pp)
      allocate(mesh_tt, source=tt)
   end subroutine

   ! illumination_p%k_in / E_in of set_illumination_parameter (main.f90:245-297)
   subroutine lib_sie_ml_fmm_set_illumination(k_in, e_in)
      real(8), intent(in) :: k_in(3), e_in(3)
      illum_k = k_in
      illum_e = e_in
      illum_set = .true.
   end subroutine

   ! lib_sie_ml_fmm_precalculation_tri (src/lib_sie_tri_mlfmm_interface.f90:59-92).
   ! Called without arguments as the reference's is, it takes the mesh of
   ! lib_sie_ml_fmm_set_mesh_files / _set_mesh; pp(3, Np) node coordinates
   ! and tt(3, Nt) 1-based node indices may also be passed directly.
   subroutine lib_sie_ml_fmm_precalculation_tri(pp, tt)
      real(8), intent(in), optional :: pp(:, :)
      integer, intent(in), optional :: tt(:, :)
      integer(c_int64_t) :: n
      if (present(pp) .and. present(tt)) then
         call spk_check(spk_setup_tri(handle, int(size(pp, 2), c_int), pp, &
                                      int(size(tt, 2), c_int), tt), 'spk_setup_tri')
      else
         if (.not. (allocated(mesh_pp) .and. allocated(mesh_tt))) then
            write(*, *) 'lib_sie_ml_fmm_precalculation_tri: no mesh (lib_sie_ml_fmm_set_mesh_files)'
            stop 1
         end if
         call spk_check(spk_setup_tri(handle, int(size(mesh_pp, 2), c_int), mesh_pp, &
                                      int(size(mesh_tt, 2), c_int), mesh_tt), 'spk_setup_tri')
      end if
      call spk_check(spk_num_edges(handle, n), 'spk_num_edges')
      m_pairs = int(n)
   end subroutine

   integer function count_lines(fname)
      character(len=*), intent(in) :: fname
      integer :: st
      character(len=1024) :: line
      count_lines = 0
      open(unit=70, file=fname, status='old', action='read')
      do
         read(70, '(a)', iostat=st) line
         if (st /= 0) exit
         if (len_trim(line) > 0) count_lines = count_lines + 1
      end do
      close(70)
   end function

   ! lib_sie_ml_fmm_calculate_vector_b_tri (:285-319): vector_b allocated by
   ! the caller (lib_math_solver_GMRES.f90:274-299), written in place.
   subroutine lib_sie_ml_fmm_calculate_vector_b_tri(vector_x, vector_b)
      double complex, dimension(:), allocatable, intent(in) :: vector_x
      double complex, dimension(:), allocatable, intent(inout) :: vector_b
      if (.not. allocated(vector_b)) allocate(vector_b(2*m_pairs))
      call spk_check(spk_matvec(handle, vector_x, vector_b), 'spk_matvec')
   end subroutine

   ! lib_sie_ml_fmm_get_vector_b_tri (:263-276) for a plane wave: called with
   ! vector_b only, as the reference's is, it takes the wave of
   ! lib_sie_ml_fmm_set_illumination; k_in, e_in may also be passed
   subroutine lib_sie_ml_fmm_get_vector_b_tri(vector_b, k_in, e_in)
      double complex, dimension(:), allocatable, intent(inout) :: vector_b
      real(8), intent(in), optional :: k_in(3), e_in(3)
      real(8) :: k(3), e(3)
      if (present(k_in) .and. present(e_in)) then
         k = k_in
         e = e_in
      else
         if (.not. illum_set) then
            write(*, *) 'lib_sie_ml_fmm_get_vector_b_tri: no illumination (lib_sie_ml_fmm_set_illumination)'
            stop 1
         end if
         k = illum_k
         e = illum_e
      end if
      if (allocated(vector_b)) deallocate(vector_b)
      allocate(vector_b(2*m_pairs))
      call spk_check(spk_rhs_plane_wave(handle, k, e, vector_b), 'spk_rhs_plane_wave')
   end subroutine

   ! lib_sie_ml_fmm_get_init_vector_x_tri (:190-239), x_initial constant
   subroutine lib_sie_ml_fmm_get_init_vector_x_tri(vector_x, x_initial)
      double complex, dimension(:), allocatable, intent(inout) :: vector_x
      real(8), intent(in) :: x_initial
      if (allocated(vector_x)) deallocate(vector_x)
      allocate(vector_x(2*m_pairs))
      vector_x = x_initial
   end subroutine

   ! lib_math_solver_gmres_run with the GPU matvec and orthogonalisation;
   ! near_field (optional): the near-field preconditioned flexible GMRES
   ! (iterations then counts its outer steps)
   subroutine lib_sie_ml_fmm_gmres_run(vector_b, vector_x, max_iterations, restart, &
                                       tolerance, orthogonalization, iterations, backward_error, &
                                       near_field)
      double complex, dimension(:), intent(in) :: vector_b
      double complex, dimension(:), intent(inout) :: vector_x
      integer, intent(in) :: max_iterations, restart, orthogonalization
      real(8), intent(in) :: tolerance
      integer, intent(out) :: iterations
      real(8), intent(out) :: backward_error
      logical, intent(in), optional :: near_field
      type(spk_gmres_params) :: p
      type(spk_gmres_info) :: info
      p%max_iterations = max_iterations
      p%restart = restart
      p%tolerance = tolerance
      p%orthogonalization = orthogonalization
      p%use_initial_guess = 1
      p%refinement_steps = 0
      p%refinement_tolerance = 0d0
      p%preconditioning = 0
      if (present(near_field)) then
         if (near_field) p%preconditioning = 2   ! SPK_PRECOND_NEAR_FIELD
      end if
      p%inner_iterations = 0
      call spk_check(spk_gmres(handle, vector_b, vector_x, p, info), 'spk_gmres')
      iterations = info%iterations
      backward_error = info%backward_error
   end subroutine

   subroutine lib_sie_ml_fmm_gpu_finalize()
      if (c_associated(handle)) call spk_destroy(handle)
      handle = c_null_ptr
   end subroutine
end module lib_sie_tri_mlfmm_gpu
