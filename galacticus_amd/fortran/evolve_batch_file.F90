!! evolve_batch_file.F90 -- Fortran host program: reads a batch of node ODE
!! systems from a binary stream file, evolves it on the GPU through the
!! Galacticus_AMD_ODE binding and writes the result.
!!
!! Input (stream, native endianness): int32 n, nT, nZ, flags(reuse, nonneg);
!! float64 tolAbs, tolRel, lnTMin, dLnT, lnZMin, dLnZ; params(48);
!! table(nT*nZ); y(n*30); nodePar(n*5); tStart(n); tEnd(n); timeStep(n).
!! Output: y(n*30), timeStep(n), scale(n*30), int32 status(n), int32 return code.
program evolve_batch_file
  use :: Galacticus_AMD_ODE, only : nodeEvolverBatched, galVariableCount, galParameterCount, galNodeParameterCount
  implicit none
  type            (nodeEvolverBatched)                :: evolver
  integer                                             :: n, countT, countZ, reuse, nonNegative, status, unitIn, unitOut
  double precision                                    :: lnTMinimum, lnTStep, lnZMinimum, lnZStep
  double precision, allocatable, dimension(:  )       :: parameters, table, timeStart, timeEnd, timeStep
  double precision, allocatable, dimension(:,:)       :: y, nodeParameters, scale
  integer         , allocatable, dimension(:  )       :: statusNode
  character(len=1024)                                 :: fileIn, fileOut

  call get_command_argument(1,fileIn )
  call get_command_argument(2,fileOut)
  open(newunit=unitIn,file=trim(fileIn),access='stream',form='unformatted',status='old',action='read')
  read(unitIn) n, countT, countZ, reuse, nonNegative
  read(unitIn) evolver%odeToleranceAbsolute, evolver%odeToleranceRelative, lnTMinimum, lnTStep, lnZMinimum, lnZStep
  allocate(parameters(galParameterCount),table(countT*countZ))
  allocate(y(n,galVariableCount),nodeParameters(n,galNodeParameterCount),scale(n,galVariableCount))
  allocate(timeStart(n),timeEnd(n),timeStep(n),statusNode(n))
  read(unitIn) parameters, table, y, nodeParameters, timeStart, timeEnd, timeStep
  close(unitIn)
  evolver%reuseODEStepSize    =reuse       /= 0
  evolver%enforceNonNegativity=nonNegative /= 0
  status=evolver%initialize(n,parameters)
  if (status == 0) status=evolver%setCoolingTable(table,countT,lnTMinimum,lnTStep,countZ,lnZMinimum,lnZStep)
  scale=0.0d0
  statusNode=0
  if (status == 0) status=evolver%evolve(y,nodeParameters,timeStart,timeEnd,timeStep,scale,statusNode)
  open(newunit=unitOut,file=trim(fileOut),access='stream',form='unformatted',status='replace',action='write')
  write(unitOut) y, timeStep, scale, statusNode, status
  close(unitOut)
  if (evolver%destroy() /= 0 .or. status /= 0) stop 1
end program evolve_batch_file
