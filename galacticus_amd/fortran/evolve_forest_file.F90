!! evolve_forest_file.F90 -- Fortran host program for whole forests: builds a
!! Parkinson-Cole-Helly forest, records the reference's walk as a schedule,
!! keeps the forest state in device memory and runs the schedule wave by wave
!! on the GPU, all through the ISO_C_BINDING modules Galacticus_AMD_ODE and
!! Galacticus_AMD_Trees.
!!
!! Input (stream, native endianness): int32 countTrees, nT, nZ; int64 seed;
!! float64 tolAbs, tolRel, lnTMin, dLnT, lnZMin, dLnZ; params(48);
!! table(nT*nZ); massRoot(countTrees); timeRoot; cosmology(7); pch(8);
!! walk(11); metalFraction, blackHoleSeed, majorRatio.
!! Output: int64 countNodes; int64 treeOffset(countTrees+1);
!! float64 y(countNodes*30); int32 status(countNodes); int32 return code.
program evolve_forest_file
  use, intrinsic :: ISO_C_Binding, only : c_int, c_int32_t, c_int64_t, c_double, c_ptr, c_null_ptr, c_loc, c_size_t
  use :: Galacticus_AMD_ODE  , only : nodeEvolverBatched, galVariableCount, galParameterCount
  use :: Galacticus_AMD_Trees
  implicit none
  type            (nodeEvolverBatched)                      :: evolver
  type            (galCosmology      )                      :: cosmology
  type            (galPCHParameters  )                      :: pch
  type            (galWalkParameters )                      :: walk
  type            (c_ptr             )                      :: forest=c_null_ptr, schedule=c_null_ptr
  type            (c_ptr             )                      :: yDevice=c_null_ptr, scaleDevice=c_null_ptr, statusDevice=c_null_ptr
  integer         (c_int             )                      :: countTrees, countT, countZ, countWaves, status, statusHIP
  integer         (c_int64_t         )                      :: seed, countNodes, countOperations
  integer                                                   :: unitIn, unitOut
  double precision                                          :: lnTMinimum, lnTStep, lnZMinimum, lnZStep, timeRoot
  double precision                                          :: metalFraction, blackHoleSeed, majorRatio
  double precision, allocatable, dimension(:  )             :: parameters, table, massRoot, mass, time
  double precision, allocatable, dimension(:,:), target     :: y
  integer         (c_int32_t), allocatable, dimension(:), target :: statusNode
  integer         (c_int32_t), allocatable, dimension(:)    :: parent, firstChild, sibling
  integer         (c_int64_t), allocatable, dimension(:)    :: treeOffset
  integer         (c_size_t )                               :: sizeState, sizeStatus
  character(len=1024)                                       :: fileIn, fileOut

  call get_command_argument(1,fileIn )
  call get_command_argument(2,fileOut)
  open(newunit=unitIn,file=trim(fileIn),access='stream',form='unformatted',status='old',action='read')
  read(unitIn) countTrees, countT, countZ
  read(unitIn) seed
  read(unitIn) evolver%odeToleranceAbsolute, evolver%odeToleranceRelative, lnTMinimum, lnTStep, lnZMinimum, lnZStep
  allocate(parameters(galParameterCount),table(countT*countZ),massRoot(countTrees))
  read(unitIn) parameters, table, massRoot, timeRoot, cosmology, pch, walk, metalFraction, blackHoleSeed, majorRatio
  close(unitIn)

  ! the forest and the walk, on the host
  status=gal_forest_build_pch(cosmology,pch,countTrees,massRoot,timeRoot,seed,forest)
  if (status == 0) status=gal_forest_size(forest,countNodes,countTrees)
  if (status == 0) status=gal_schedule_build(forest,cosmology,walk,schedule)
  if (status == 0) status=gal_schedule_size(schedule,countOperations,countWaves)
  if (status /= 0) stop 2
  allocate(y(countNodes,galVariableCount),statusNode(countNodes),treeOffset(countTrees+1))
  allocate(parent(countNodes),firstChild(countNodes),sibling(countNodes),mass(countNodes),time(countNodes))
  status=gal_forest_arrays(forest,treeOffset,parent,firstChild,sibling,mass,time)
  if (status == 0) status=gal_forest_initial_state(forest,cosmology,walk,metalFraction,blackHoleSeed,int(countNodes,c_int),y)

  ! the node evolver and the forest state in device memory
  evolver%reuseODEStepSize=.true.
  if (status == 0) status=evolver%initialize(int(countNodes),parameters)
  if (status == 0) status=evolver%setCoolingTable(table,countT,lnTMinimum,lnTStep,countZ,lnZMinimum,lnZStep)
  sizeState =int(8,c_size_t)*int(galVariableCount,c_size_t)*int(countNodes,c_size_t)
  sizeStatus=int(4,c_size_t)*int(countNodes,c_size_t)
  statusHIP=hipMalloc(yDevice,sizeState)+hipMalloc(scaleDevice,sizeState)+hipMalloc(statusDevice,sizeStatus)
  if (statusHIP /= 0) stop 3
  statusHIP=hipMemcpy(yDevice,c_loc(y),sizeState,hipMemcpyHostToDevice)

  ! every wave of the schedule on the GPU
  if (status == 0 .and. statusHIP == 0) &
       & status=gal_tree_evolve_d(evolver%handle,schedule,int(countNodes,c_int),yDevice,scaleDevice,statusDevice, &
       &                          majorRatio,c_null_ptr)
  statusHIP=hipMemcpy(c_loc(y         ),yDevice     ,sizeState ,hipMemcpyDeviceToHost) &
       &   +hipMemcpy(c_loc(statusNode),statusDevice,sizeStatus,hipMemcpyDeviceToHost)
  if (statusHIP /= 0 .and. status == 0) status=-2

  open(newunit=unitOut,file=trim(fileOut),access='stream',form='unformatted',status='replace',action='write')
  write(unitOut) countNodes, treeOffset, y, statusNode, status
  close(unitOut)
  statusHIP=hipFree(yDevice)+hipFree(scaleDevice)+hipFree(statusDevice)
  if (gal_schedule_destroy(schedule) /= 0 .or. gal_forest_destroy(forest) /= 0) stop 4
  if (evolver%destroy() /= 0 .or. status /= 0) stop 1
end program evolve_forest_file
