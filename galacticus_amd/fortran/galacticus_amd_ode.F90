!! galacticus_amd_ode.F90 -- Fortran host binding of the MI355X node-ODE engine.
!!
!! ISO_C_BINDING interfaces to the C ABI of include/gal_ode.h, and a node
!! evolver type that keeps the parameter keys of the reference's
!! mergerTreeNodeEvolverStandard (source/merger_trees/node_evolver/standard.F90:165-256)
!! and its isAccurate method (standard.F90:757-768, Values_Agree of
!! source/numerical/comparison.F90:129-155).  The reference reaches its ODE
!! solver through bind(c) interfaces of the same style
!! (source/numerical/ODE_solver/solver.F90:62-190).
!!
!! The type carries the method set of the reference's mergerTreeNodeEvolver
!! class (source/merger_trees/node_evolver/_class.F90:41-79) for a batch of
!! nodes: evolve (with per-node status and an optional systemClockMaximum,
!! past which nodes return errorStatusXCPU), promote, merge and isAccurate;
!! readParameters takes odeToleranceAbsolute, odeToleranceRelative,
!! reuseODEStepSize and enforceNonNegativity from the mergerTreeNodeEvolver
!! element of a Galacticus parameter file, with the reference's defaults
!! (standard.F90:165-256).
module Galacticus_AMD_ODE
  use, intrinsic :: ISO_C_Binding, only : c_int, c_double, c_ptr, c_null_ptr, c_associated, c_int64_t, c_loc
  implicit none
  private
  public :: nodeEvolverBatched, galVariableCount, galParameterCount, galNodeParameterCount, &
       &    errorStatusSuccess, errorStatusUnderflow, errorStatusXCPU

  integer(c_int), parameter :: galVariableCount     =30
  integer(c_int), parameter :: galParameterCount    =48
  integer(c_int), parameter :: galNodeParameterCount= 7
  integer(c_int), parameter :: flagEnforceNonNegativity=1, flagReuseStepSize=2, flagCostOrder=8
  integer(c_int), parameter :: eventPromote=1, eventMerge=3
  integer       , parameter :: errorStatusSuccess=0, errorStatusUnderflow=15, errorStatusXCPU=1025

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
     function gal_ode_batch_set_wall_limit(handle,seconds) bind(c,name='gal_ode_batch_set_wall_limit')
       import c_int, c_double, c_ptr
       type   (c_ptr   ), value :: handle
       real   (c_double), value :: seconds
       integer(c_int   )        :: gal_ode_batch_set_wall_limit
     end function gal_ode_batch_set_wall_limit
     function gal_node_events_h(countEvents,kind,node,target,countNodes,y,timeStep,ratioMajor) bind(c,name='gal_node_events_h')
       import c_int, c_double
       integer(c_int   ), value        :: countEvents, countNodes
       integer(c_int   ), dimension(*) :: kind, node, target
       real   (c_double), dimension(*) :: y, timeStep
       real   (c_double), value        :: ratioMajor
       integer(c_int   )               :: gal_node_events_h
     end function gal_node_events_h
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
     procedure :: readParameters  => batchedReadParameters
     procedure :: initialize      => batchedInitialize
     procedure :: setCoolingTable => batchedSetCoolingTable
     procedure :: evolve          => batchedEvolve
     procedure :: promote         => batchedPromote
     procedure :: merge           => batchedMerge
     procedure :: isAccurate      => batchedIsAccurate
     procedure :: destroy         => batchedDestroy
  end type nodeEvolverBatched

contains

  integer function batchedReadParameters(self,fileName) result(status)
    !! Read the node evolver's parameters from the mergerTreeNodeEvolver
    !! element of a parameter file; absent keys keep the reference's defaults.
    !! Status: 0, or 1 if the file cannot be read.
    class    (nodeEvolverBatched), intent(inout) :: self
    character(len=*             ), intent(in   ) :: fileName
    character(len=:             ), allocatable   :: text, block
    character(len=64            )                :: value
    integer                                      :: unit, length, ioStatus, i, j
    logical                                      :: found
    self%odeToleranceAbsolute=0.01d0
    self%odeToleranceRelative=0.01d0
    self%reuseODEStepSize    =.true.
    self%enforceNonNegativity=.false.
    status=1
    inquire(file=fileName,size=length)
    if (length <= 0) return
    allocate(character(len=length) :: text)
    open(newunit=unit,file=fileName,access='stream',form='unformatted',status='old',action='read',iostat=ioStatus)
    if (ioStatus /= 0) return
    read(unit,iostat=ioStatus) text
    close(unit)
    if (ioStatus /= 0) return
    status=0
    ! drop comments
    do
       i=index(text,'<!--')
       if (i == 0) exit
       j=index(text(i+4:),'-->')
       if (j == 0) then
          text=text(1:i-1)
       else
          text=text(1:i-1)//text(i+4+j+2:)
       end if
    end do
    i=elementStart(text,'mergerTreeNodeEvolver')
    if (i == 0) return
    j=index(text(i:),'</mergerTreeNodeEvolver>')
    if (j == 0) return
    block=text(i:i+j-1)
    call attributeValue(block,'odeToleranceAbsolute',value,found)
    if (found) read(value,*) self%odeToleranceAbsolute
    call attributeValue(block,'odeToleranceRelative',value,found)
    if (found) read(value,*) self%odeToleranceRelative
    call attributeValue(block,'reuseODEStepSize'    ,value,found)
    if (found) self%reuseODEStepSize    =trim(adjustl(value)) == 'true'
    call attributeValue(block,'enforceNonNegativity',value,found)
    if (found) self%enforceNonNegativity=trim(adjustl(value)) == 'true'
  end function batchedReadParameters

  integer function elementStart(text,name) result(position)
    !! Position of the opening tag <name ...> (name followed by a blank, / or >), or 0.
    character(len=*), intent(in) :: text, name
    integer                      :: offset, i, k
    character                    :: next
    position=0
    offset  =0
    do
       i=index(text(offset+1:),'<'//name)
       if (i == 0) return
       k=offset+i+len(name)+1
       if (k <= len(text)) then
          next=text(k:k)
          if (next == ' ' .or. next == '/' .or. next == '>' .or. next == char(9) .or. next == char(10) &
               & .or. next == char(13)) then
             position=offset+i
             return
          end if
       end if
       offset=offset+i
    end do
  end function elementStart

  subroutine attributeValue(text,name,value,found)
    !! The value="..." attribute of the first <name .../> element in text.
    character(len=*), intent(in   ) :: text, name
    character(len=*), intent(  out) :: value
    logical         , intent(  out) :: found
    integer                         :: i, j, k
    value=''
    found=.false.
    i=elementStart(text,name)
    if (i == 0) return
    j=index(text(i:),'>')
    if (j == 0) return
    k=index(text(i:i+j-1),'value="')
    if (k == 0) return
    k=i+k-1+len('value="')
    j=index(text(k:),'"')
    if (j == 0) return
    value=text(k:k+j-2)
    found=.true.
  end subroutine attributeValue

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

  integer function batchedEvolve(self,y,nodeParameters,timeStart,timeEnd,timeStep,scale,statusNode,systemClockMaximum) &
       & result(status)
    !! Evolve all nodes; y(n,30), nodeParameters(n,7) (column-major = the engine's structure of arrays).
    !! With systemClockMaximum (a System_Clock count), nodes not finished when
    !! the count is passed return errorStatusXCPU with their state as it came
    !! in (node_evolver/standard.F90:693-705); the limit is checked on the GPU.
    class           (nodeEvolverBatched), intent(inout)                 :: self
    integer         (kind=8            ), intent(in   ), optional       :: systemClockMaximum
    double precision                    , intent(inout), dimension(:,:) :: y
    double precision                    , intent(in   ), dimension(:,:) :: nodeParameters
    double precision                    , intent(in   ), dimension(:  ) :: timeStart, timeEnd
    double precision                    , intent(inout), dimension(:  ) :: timeStep
    double precision                    , intent(inout), dimension(:,:) :: scale
    integer                             , intent(  out), dimension(:  ) :: statusNode
    integer(c_int), allocatable, dimension(:) :: status_
    double precision, allocatable, dimension(:,:) :: parameters_
    double precision, allocatable, dimension(:  ) :: start_, end_
    integer          :: n
    integer(kind=8)  :: clockCount, clockRate
    double precision :: seconds
    n=size(timeStart)
    allocate(status_(n))
    parameters_=nodeParameters
    start_     =timeStart
    end_       =timeEnd
    if (present(systemClockMaximum)) then
       if (systemClockMaximum > 0_8) then
          call System_Clock(clockCount,clockRate)
          ! already past the limit: the shortest limit the device can check
          seconds=max(dble(systemClockMaximum-clockCount)/dble(clockRate),1.0d-9)
          status=gal_ode_batch_set_wall_limit(self%handle,seconds)
          if (status /= 0) return
       end if
    end if
    status=gal_ode_batch_evolve_h(self%handle,int(n,c_int),y,parameters_,start_,end_,timeStep,scale,status_,c_null_ptr)
    statusNode=status_
    if (present(systemClockMaximum)) then
       if (gal_ode_batch_set_wall_limit(self%handle,0.0d0) /= 0 .and. status == 0) status=-1
    end if
  end function batchedEvolve

  integer function batchedPromote(self,y,timeStep,node,parent) result(status)
    !! Promote node into its parent (node_evolver/standard.F90:1241-1356): the
    !! parent takes the node's galaxy state and step size.  Indices are 1-based.
    class           (nodeEvolverBatched), intent(inout)                 :: self
    double precision                    , intent(inout), dimension(:,:) :: y
    double precision                    , intent(inout), dimension(:  ) :: timeStep
    integer                             , intent(in   )                 :: node, parent
    integer(c_int), dimension(1) :: kind, from, to
    kind=eventPromote
    from=int(node  -1,c_int)
    to  =int(parent-1,c_int)
    status=gal_node_events_h(1_c_int,kind,from,to,int(size(timeStep),c_int),y,timeStep,0.25d0)
  end function batchedPromote

  integer function batchedMerge(self,y,timeStep,satellite,host,ratioMajor) result(status)
    !! Merge satellite into host (mergerMassMovementsSimple,
    !! satellites/merging/mass_movements/simple.F90:208-290; major merger when
    !! the smaller galaxy is at least ratioMajor of the larger, default 0.25).
    !! Indices are 1-based.
    class           (nodeEvolverBatched), intent(inout)                 :: self
    double precision                    , intent(inout), dimension(:,:) :: y
    double precision                    , intent(inout), dimension(:  ) :: timeStep
    integer                             , intent(in   )                 :: satellite, host
    double precision                    , intent(in   ), optional       :: ratioMajor
    integer(c_int), dimension(1) :: kind, from, to
    double precision             :: ratio
    ratio=0.25d0
    if (present(ratioMajor)) ratio=ratioMajor
    kind=eventMerge
    from=int(satellite-1,c_int)
    to  =int(host     -1,c_int)
    status=gal_node_events_h(1_c_int,kind,from,to,int(size(timeStep),c_int),y,timeStep,ratio)
  end function batchedMerge

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
