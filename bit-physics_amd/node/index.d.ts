// TypeScript declarations of the D3Q19 Node-API addon (libd3q19 over its C ABI).
// Populations are (19, Nx, Ny, Nz) z-fastest in the reference's lexicographic
// velocity order; vector fields are (3, Nx, Ny, Nz).

export type Storage = 'f64' | 'f32' | 'f16' | 'bf16' | 'bf16sr' | 'me16' | 'me16ref';
export type Boundary = 'periodic' | 'ywalls' | 'halfway';
export type Face = '-x' | '+x' | '-y' | '+y' | '-z' | '+z';
export type Vec3 = [number, number, number];

export interface LatticeConfig {
  nx: number;
  ny: number;
  nz: number;
  tau: number;
  storage?: Storage;            // default 'f32'
  boundary?: Boundary;          // default 'periodic'
  wallVelocityTop?: Vec3;       // 'ywalls': apply_bounce_back_y_walls(wall_velocity_top)
  wallVelocityBottom?: Vec3;    // 'ywalls': apply_bounce_back_y_walls(wall_velocity_bottom)
  wallFaces?: Face[];           // 'halfway': wall faces (others periodic)
  faceVelocity?: Partial<Record<Face, Vec3>>;
  force?: Vec3;                 // uniform Guo body force
  device?: number;
}

export interface Moments {
  rho: Float64Array;            // (Nx, Ny, Nz)
  u: Float64Array;              // (3, Nx, Ny, Nz), Guo-corrected
}

export interface Diagnostics {
  mass: number;
  momentum: Vec3;
  kineticEnergy: number;
}

export interface Health {
  checkHealth: number;          // 0 when every population, rho and u is finite, else 1
  nonfiniteCells: number;
  rhoMin: number;
  rhoMax: number;
  uMaxLat: number;              // max |u| over the finite cells
}

export declare class Lattice {
  constructor(cfg: LatticeConfig);
  readonly shape: Vec3;
  setEquilibrium(rho: Float64Array | Float32Array | null, u: Float64Array | Float32Array): void;
  setPopulations(f: Float64Array | Float32Array): void;
  step(n?: number): void;
  readPopulations(dtype?: 'f64' | 'f32'): Float64Array | Float32Array;
  readMoments(): Moments;
  diagnostics(): Diagnostics;
  stepCount(): number;
  // 'me16' / 'me16ref' storage: quantization ranges of the 19 moments of f, before the state is loaded
  setMomentRanges(lo: Float64Array, hi: Float64Array): void;
  // minimum and maximum over all cells of the 19 moments of the current state
  // stored: of the state as the lattice stores it (post-collision, what 'me16' encodes)
  momentRange(stored?: boolean): { lo: Float64Array; hi: Float64Array };
  health(): Health;
  close(): void;
}

export declare function version(): string;
