"""MI355X-native curvilinear velocity-stress solver (2D P-SV elastic / Qp-Qs GSLS, 3D nine-field).

Host-side mirror of the reference's session interface; the compute path is the HIP
library ``csrc/libswl.so`` reached through the C-ABI declared in ``include/swl.h``.
"""
from .params import (Absorber, AbsorbingBoundarySettings3D, BackendKind, CfsPmlSettings, GslsSettings, Layer,
                     LayerDepthMode3D, MaterialLayer3D, Parameters, Precision, ReceiverArrayKind3D,
                     ReceiverArraySettings3D, SimulationParameters, SourceKind, SourceMechanism, SurfaceGrid3D)
from .simulation import (AsyncSimulation, AsyncSimulation3D, Batch2D, Receiver, ReceiverTrace3D, Simulation, Simulation3D,
                         estimateModelMemory3D, memoryBudget, writeReceivers)
from . import modelio, native, records, workloads
from .records import ImportedRecords, SegyExportProfile, TraceRecord, read_segy, write_segy
from .modelio import (ComputeSettings, ModelDocument, SnapshotExportOptions, VtsExportOptions3D,
                      displayedSnapshotOptions, readModelVts, writeModelVts, writeReceiversCsv, writeSnapshot, writeVts)
