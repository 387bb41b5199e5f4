// HipPhysicsBackend.cs -- the reference-side binding of libcmhip.so.
//
// Drop-in IPhysicsComputeBackend + IGravityModelAwareBackend for
// src/CelestialMechanics.Physics (mfscpayload-690/celestial-mechanics-desktop),
// in the slot CudaPhysicsBackend reserves (Solvers/CudaPhysicsBackend.cs:35-48).
// Not compiled in this repository (no .NET toolchain here); it is the file a
// maintainer adds to src/CelestialMechanics.Physics/Solvers/.  Requires
// <AllowUnsafeBlocks>true</AllowUnsafeBlocks> in the project.
using System;
using System.Runtime.InteropServices;
using CelestialMechanics.Physics.SoA;

namespace CelestialMechanics.Physics.Solvers;

internal static unsafe class CmHip
{
    private const string Lib = "cmhip";   // libcmhip.so / cmhip.dll

    [StructLayout(LayoutKind.Sequential)]
    public struct BackendConfig
    {
        public int Kind;            // 0 pairwise (CpuSingleThreadBackend), 1 sweep (CpuParallelBackend), 2 Barnes-Hut
        public int Precision;       // 0 FP64, 1 FP64 fast rsqrt, 2 FP32 split
        public int ShellTheorem;
        public int MortonLevels;    // 10 -> 30-bit keys
        public double Theta;
        public int LeafCapacity;
        public int GroupSize;
        public int LearnedCorrection; // 1: MLP far-field correction (precision 2 only)
        public double FarFieldRatio;  // 0 off; kappa <= 0.5: group Taylor expansion of far cells (precision 2)
    }

    [StructLayout(LayoutKind.Sequential)]
    public struct SoAPtrs
    {
        public double* PosX, PosY, PosZ, VelX, VelY, VelZ, AccX, AccY, AccZ;
        public double* OldAccX, OldAccY, OldAccZ, Mass, Radius, Density;
        public byte* IsActive, IsCollidable;   // bool[] travels pinned as byte*
        public int* BodyTypeIndex;
    }

    [StructLayout(LayoutKind.Sequential)]
    public struct PnConfig
    {
        public int Enabled;
        public double MaxVelocityFractionC;        // 0.3
        public double SchwarzschildWarningFactor;  // 3.0
    }

    [StructLayout(LayoutKind.Sequential)]
    public struct PnWarning
    {
        public int Body, Other;
        public double Distance, SchwarzschildRadius;
    }

    [StructLayout(LayoutKind.Sequential)]
    public struct SofteningConfig
    {
        public int Mode;                          // 0 global, 1 per body type, 2 adaptive
        public fixed double TypeSoftening[16];
        public double AdaptiveScale;              // 0.01
        public double AdaptiveMin;                // 1e-8
    }

    [StructLayout(LayoutKind.Sequential)]
    public struct CollisionConfig
    {
        public int Mode;                          // CollisionMode: 0 MergeOnly, 1 BounceOnly, 2 Realistic
        public double DefaultRestitution;
        public double FragmentationSpecificEnergyThreshold;
        public double FragmentationMassLossCap;
        public double CaptureVelocityFactor;
        public int PromoteCompactRemnants;
        public int MaxMergesPerFrame;             // <= 0: no cap
        public int MaxMergesPerBody;              // <= 0: no cap
    }

    [DllImport(Lib)] public static extern int cm_create(int capacity, int device, out IntPtr engine);
    [DllImport(Lib)] public static extern int cm_destroy(IntPtr engine);
    [DllImport(Lib)] public static extern IntPtr cm_last_error(IntPtr engine);
    [DllImport(Lib)] public static extern int cm_default_config(ref BackendConfig cfg);
    [DllImport(Lib)] public static extern int cm_configure(IntPtr engine, ref BackendConfig cfg);
    [DllImport(Lib)] public static extern int cm_upload_soa(IntPtr engine, ref SoAPtrs src, int count);
    [DllImport(Lib)] public static extern int cm_download_soa(IntPtr engine, ref SoAPtrs dst, int count);
    [DllImport(Lib)] public static extern int cm_upload_force_inputs(IntPtr engine, ref SoAPtrs src, int count);
    [DllImport(Lib)] public static extern int cm_compute_forces(IntPtr engine, double softening);
    [DllImport(Lib)] public static extern int cm_verlet_step(IntPtr engine, double softening, double dt, int nSteps);
    [DllImport(Lib)] public static extern int cm_yoshida4_step(IntPtr engine, double softening, double dt, int nSteps);
    [DllImport(Lib)] public static extern int cm_default_pn_config(ref PnConfig cfg);
    [DllImport(Lib)] public static extern int cm_set_post_newtonian(IntPtr engine, ref PnConfig cfg);
    [DllImport(Lib)] public static extern int cm_pn_warnings(IntPtr engine, PnWarning* outp, int cap, out int nTotal);
    [DllImport(Lib)] public static extern int cm_default_softening_config(ref SofteningConfig cfg);
    [DllImport(Lib)] public static extern int cm_set_softening(IntPtr engine, ref SofteningConfig cfg);
    [DllImport(Lib)] public static extern int cm_default_collision_config(ref CollisionConfig cfg);
    [DllImport(Lib)] public static extern int cm_configure_collisions(IntPtr engine, ref CollisionConfig cfg);
    // cm_energy_report: EnergyCalculator quantities plus the centre of mass of
    // the engine's GPU energy snapshot.
    [StructLayout(LayoutKind.Sequential)]
    public struct EnergyReport
    {
        public double Kinetic;
        public double Potential;
        public fixed double Momentum[3];
        public fixed double AngularMomentum[3];
        public double AngularMomentumMag;
        public double TotalMass;
        public int ActiveCount;
        public fixed double CenterOfMass[3];
    }

    [DllImport(Lib)] public static extern int cm_resolve_collisions(IntPtr engine, out int nResolved);
    // withPotential: 0 none, 1 exact pair sum, 2 Barnes-Hut potential
    [DllImport(Lib)] public static extern int cm_energy(IntPtr engine, double softening, int withPotential, out EnergyReport report);
    [DllImport(Lib)] public static extern int cm_max_acceleration(IntPtr engine, out double maxAccel);
    // PlummerSphere.Generate(count, totalMass, scaleLength, seed) body for body
    [DllImport(Lib)] public static extern int cm_generate_plummer_dotnet(int n, double totalMass, double scaleLength, int seed, ref SoAPtrs output);

    public static void Check(IntPtr engine, int rc, string what)
    {
        if (rc == 0) return;
        string msg = Marshal.PtrToStringAnsi(cm_last_error(engine)) ?? "";
        throw new InvalidOperationException($"{what} failed ({rc}): {msg}");
    }
}

/// <summary>
/// GPU (MI355X) backend: uploads the SoA state, computes accelerations on the
/// device (Barnes-Hut by default) and writes AccX/Y/Z for every index below
/// Count, inactive bodies getting 0 -- the IPhysicsComputeBackend contract.
/// </summary>
public sealed unsafe class HipPhysicsBackend : IPhysicsComputeBackend, IGravityModelAwareBackend, IDisposable
{
    private IntPtr _engine;
    private CmHip.BackendConfig _cfg;
    private readonly int _capacity;

    public HipPhysicsBackend(int capacity, double theta = 0.5, int precision = 2)
    {
        _capacity = capacity;
        CmHip.Check(IntPtr.Zero, CmHip.cm_create(capacity, 0, out _engine), "cm_create");
        CmHip.cm_default_config(ref _cfg);
        _cfg.Theta = theta;
        _cfg.Precision = precision;
        CmHip.Check(_engine, CmHip.cm_configure(_engine, ref _cfg), "cm_configure");
    }

    public bool EnableShellTheorem
    {
        get => _cfg.ShellTheorem != 0;
        set { _cfg.ShellTheorem = value ? 1 : 0; CmHip.Check(_engine, CmHip.cm_configure(_engine, ref _cfg), "cm_configure"); }
    }

    public double Theta
    {
        get => _cfg.Theta;
        set { _cfg.Theta = value; CmHip.Check(_engine, CmHip.cm_configure(_engine, ref _cfg), "cm_configure"); }
    }

    /// <summary>
    /// PostNewtonianBackend around this backend (Solvers/PostNewtonianBackend.cs:33-43):
    /// every force evaluation adds PostNewtonian1Correction's term on the device.
    /// </summary>
    public bool EnablePostNewtonian
    {
        get => _pn.Enabled != 0;
        set
        {
            if (_pn.MaxVelocityFractionC == 0.0) CmHip.cm_default_pn_config(ref _pn);
            _pn.Enabled = value ? 1 : 0;
            CmHip.Check(_engine, CmHip.cm_set_post_newtonian(_engine, ref _pn), "cm_set_post_newtonian");
        }
    }

    /// <summary>PostNewtonian1Correction.OnSchwarzschildProximityWarning: raised after
    /// ComputeForces for every strong-field pair the device skipped, in (i, j) order.</summary>
    public Action<int, int, double, double>? OnSchwarzschildProximityWarning { get; set; }

    private CmHip.PnConfig _pn;

    // allocated once: ComputeForces allocates nothing per call
    private readonly CmHip.PnWarning[] _warnings = new CmHip.PnWarning[4096];

    private void RaiseProximityWarnings()
    {
        if (_pn.Enabled == 0 || OnSchwarzschildProximityWarning == null) return;
        var buf = _warnings;
        fixed (CmHip.PnWarning* w = buf)
        {
            CmHip.Check(_engine, CmHip.cm_pn_warnings(_engine, w, buf.Length, out int total), "cm_pn_warnings");
            for (int k = 0; k < Math.Min(total, buf.Length); k++)
                OnSchwarzschildProximityWarning(buf[k].Body, buf[k].Other, buf[k].Distance, buf[k].SchwarzschildRadius);
        }
    }

    public void ComputeForces(BodySoA bodies, double softening)
    {
        if (bodies.Count > _capacity) throw new InvalidOperationException("body count exceeds backend capacity");
        // only what the force evaluation reads crosses PCIe (positions,
        // masses, radii, active flags; velocities and body types when the 1PN
        // decorator or per-type softening reads them), and only AccX/Y/Z back
        fixed (double* px = bodies.PosX, py = bodies.PosY, pz = bodies.PosZ,
               vx = bodies.VelX, vy = bodies.VelY, vz = bodies.VelZ,
               ax = bodies.AccX, ay = bodies.AccY, az = bodies.AccZ,
               m = bodies.Mass, r = bodies.Radius)
        fixed (bool* act = bodies.IsActive)
        fixed (int* type = bodies.BodyTypeIndex)
        {
            var p = new CmHip.SoAPtrs
            {
                PosX = px, PosY = py, PosZ = pz, VelX = vx, VelY = vy, VelZ = vz,
                Mass = m, Radius = r, IsActive = (byte*)act, BodyTypeIndex = type,
            };
            CmHip.Check(_engine, CmHip.cm_upload_force_inputs(_engine, ref p, bodies.Count),
                        "cm_upload_force_inputs");
            CmHip.Check(_engine, CmHip.cm_compute_forces(_engine, softening), "cm_compute_forces");
            var outp = new CmHip.SoAPtrs { AccX = ax, AccY = ay, AccZ = az };
            CmHip.Check(_engine, CmHip.cm_download_soa(_engine, ref outp, bodies.Count), "cm_download_soa");
        }
        RaiseProximityWarnings();
    }

    public void Dispose()
    {
        if (_engine != IntPtr.Zero) { CmHip.cm_destroy(_engine); _engine = IntPtr.Zero; }
    }
}
