!! galacticus_amd_ode.F90 -- Fortran host binding of the MI355X node-ODE engine.
!!
!! ISO_C_BINDING interfaces to the C ABI of include/gal_ode.h, and a node
!! evolver type that keeps the parameter keys of the reference's
!! mergerTreeNodeEvolverStandard (source/merger_trees/node_evolver/standard.F90:165-256)
!! and its isAccurate method (standard.F90:757-768, Values_Agree of
!! source/numerical/comparison.F90:129-155).  The reference reaches its ODE
!! solver through bind(c) interfaces of the same style
!! (source/numerical/ODE_solver/solver.F90:62-190).
module Galacticus_AMD_ODE
  use, intrinsic :: ISO_C_Binding, only : c_int, c_double, c_ptr, c_null_ptr, c_associated, c_int64_t, c_loc
  implicit none
  private
  public :: nodeEvolverBatched, galVariableCount, galParameterCount, galNodeParameterCount

  integer(c_int), parameter :: galVariableCount     =30
  integer(c_int), parameter :: galParameterCount    =48
  integer(c_int), parameter :: galNodeParameterCount= 7
  integer(c_int), parameter :: flagEnforceNonNegativity=1, flagReuseStepSize=2, flagCostOrder=8

  interface
     function gal_ode_batch_create(nMax,toleranceAbsolute,toleranceRelative,flags,handle) bind(c,name='gal_ode_batch_create')
       import c_int, c_double, c_ptr
       integer(c_int), value :: nMax, flags
       real   (c_double), value :: toleranceAbsolute, toleranceRelative
       type   (c_ptr   )        :: handle
       integer(c_int)           :: gal_ode_batch_create
     end function gal_ode_batch_create
     function gal_ode_batch_destroy(handle) bind(c,name='gal_ode_batch_destroy')
       import c_int, c_ptr
       type   (c_ptr), value :: handle
       integer(c_int)        :: gal_ode_batch_destroy
     end function gal_ode_batch_destroy
     function gal_ode_batch_set_physics(handle,parameters,countParameters) bind(c,name='gal_ode_batch_set_physics')
       import c_int, c_double, c_ptr
       type   (c_ptr   ), value         :: handle
       real   (c_double), dimension(*)  :: parameters
       integer(c_int   ), value         :: countParameters
       integer(c_int   )                :: gal_ode_batch_set_physics
     end function gal_ode_batch_set_physics
     function gal_ode_batch_set_cooling_table(handle,lnLambda,countT,lnTMinimum,lnTStep,countZ,lnZMinimum,lnZStep) &
          & bind(c,name='gal_ode_batch_set_cooling_table')
       import c_int, c_double, c_ptr
       type   (c_ptr   ), value        :: handle
       real   (c_double), dimension(*) :: lnLambda
       integer(c_int   ), value        :: countT, countZ
       real   (c_double), value        :: lnTMinimum, lnTStep, lnZMinimum, lnZStep
       integer(c_int   )               :: gal_ode_batch_set_cooling_table
     end function gal_ode_batch_set_cooling_table
     function gal_ode_batch_evolve_h(handle,n,y,nodeParameters,timeStart,timeEnd,timeStep,scale,status,counters) &
          & bind(c,name='gal_ode_batch_evolve_h')
       import c_int, c_double, c_ptr
       type   (c_ptr   ), value        :: handle
       integer(c_int   ), value        :: n
       real   (c_double), dimension(*) :: y, nodeParameters, timeStart, timeEnd, timeStep, scale
       integer(c_int   ), dimension(*) :: status
       type   (c_ptr   ), value        :: counters
       integer(c_int   )               :: gal_ode_batch_evolve_h
     end function gal_ode_batch_evolve_h
  end interface

  !! Batched node evolver: one call evolves n nodes, each from timeStart(i)
  !! to timeEnd(i); state arrays are structure-of-arrays y(i,v) = variable v of node i.
  type :: nodeEvolverBatched
     type            (c_ptr) :: handle                =c_null_ptr
     integer                 :: countMaximum          =0
     double precision        :: odeToleranceAbsolute  =0.01d0, odeToleranceRelative=0.01d0
     logical                 :: reuseODEStepSize      =.true. , enforceNonNegativity=.false.
   contains
     procedure :: initialize      => batchedInitialize
     procedure :: setCoolingTable => batchedSetCoolingTable
     procedure :: evolve          => batchedEvolve
     procedure :: isAccurate      => batchedIsAccurate
     procedure :: destroy         => batchedDestroy
  end type nodeEvolverBatched

contains

  integer function batchedInitialize(self,countMaximum,parameters) result(status)
    class           (nodeEvolverBatched), intent(inout)               :: self
    integer                             , intent(in   )               :: countMaximum
    double precision                    , intent(in   ), dimension(:) :: parameters
    integer(c_int)                                                    :: flags
    double precision                                   , dimension(galParameterCount) :: parameters_
    flags=flagCostOrder
    if (self%reuseODEStepSize    ) flags=ior(flags,flagReuseStepSize       )
    if (self%enforceNonNegativity) flags=ior(flags,flagEnforceNonNegativity)
    status=gal_ode_batch_create(int(countMaximum,c_int),self%odeToleranceAbsolute,self%odeToleranceRelative,flags,self%handle)
    if (status /= 0) return
    self%countMaximum=countMaximum
    parameters_=parameters
    status=gal_ode_batch_set_physics(self%handle,parameters_,galParameterCount)
  end function batchedInitialize

  integer function batchedSetCoolingTable(self,lnLambda,countT,lnTMinimum,lnTStep,countZ,lnZMinimum,lnZStep) result(status)
    class           (nodeEvolverBatched), intent(inout)               :: self
    double precision                    , intent(in   ), dimension(:) :: lnLambda
    integer                             , intent(in   )               :: countT, countZ
    double precision                    , intent(in   )               :: lnTMinimum, lnTStep, lnZMinimum, lnZStep
    double precision, allocatable, dimension(:) :: table
    table=lnLambda
    status=gal_ode_batch_set_cooling_table(self%handle,table,int(countT,c_int),lnTMinimum,lnTStep,int(countZ,c_int),lnZMinimum,lnZStep)
  end function batchedSetCoolingTable

  integer function batchedEvolve(self,y,nodeParameters,timeStart,timeEnd,timeStep,scale,statusNode) result(status)
    !! Evolve all nodes; y(n,30), nodeParameters(n,7) (column-major = the engine's structure of arrays).
    class           (nodeEvolverBatched), intent(inout)                 :: self
    double precision                    , intent(inout), dimension(:,:) :: y
    double precision                    , intent(in   ), dimension(:,:) :: nodeParameters
    double precision                    , intent(in   ), dimension(:  ) :: timeStart, timeEnd
    double precision                    , intent(inout), dimension(:  ) :: timeStep
    double precision                    , intent(inout), dimension(:,:) :: scale
    integer                             , intent(  out), dimension(:  ) :: statusNode
    integer(c_int), allocatable, dimension(:) :: status_
    double precision, allocatable, dimension(:,:) :: parameters_
    double precision, allocatable, dimension(:  ) :: start_, end_
    integer :: n
    n=size(timeStart)
    allocate(status_(n))
    parameters_=nodeParameters
    start_     =timeStart
    end_       =timeEnd
    status=gal_ode_batch_evolve_h(self%handle,int(n,c_int),y,parameters_,start_,end_,timeStep,scale,status_,c_null_ptr)
    statusNode=status_
  end function batchedEvolve

  elemental logical function batchedIsAccurate(self,valueNode,valueExpected)
    !! Values_Agree with the evolver's tolerances (numerical/comparison.F90:129-155).
    class           (nodeEvolverBatched), intent(in) :: self
    double precision                    , intent(in) :: valueNode, valueExpected
    batchedIsAccurate= abs(valueNode-valueExpected) <= self%odeToleranceAbsolute                                  &
         &            .or. abs(valueNode-valueExpected) <= 0.5d0*abs(valueNode+valueExpected)*self%odeToleranceRelative &
         &            .or. valueNode == valueExpected
  end function batchedIsAccurate

  integer function batchedDestroy(self) result(status)
    class(nodeEvolverBatched), intent(inout) :: self
    status=0
    if (c_associated(self%handle)) status=gal_ode_batch_destroy(self%handle)
    self%handle=c_null_ptr
  end function batchedDestroy

end module Galacticus_AMD_ODE
