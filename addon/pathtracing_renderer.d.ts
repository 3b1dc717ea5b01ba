// Types of the Node drop-in (mirrors packages/lib/renderer.ts PathtracingRenderer).
export interface PathtracingCamera {
  fov: number;
  near: number;
  matrixWorld: { elements: ArrayLike<number> };
  position: { x: number; y: number; z: number };
}
export interface SceneBuffers {
  vertexBuffer: Float32Array;          // 20 f32 per vertex
  triangleIndexBuffer: Uint32Array;
  bvhPositionBuffer: Float32Array;
  bvhIndexBuffer: Uint32Array;
  materials: { data: Float32Array }[]; // 68 f32 each
  lights?: { position: number[]; emission: number[] }[];
  bvhNodes?: Float32Array;             // optional prebuilt bvh_gpu nodes
  bvhTriangleIndices?: Int32Array;
  // material textures: TexInfo records (8 f32 each, material.ts:483-523), RGBA8 texture
  // arrays (width, height, layers per array) and their texels, row 0 first
  textures?: { texInfos: Float32Array; arrayDims: Int32Array; texels: Uint8Array };
  // two-level scene (RenderServer.addScene only): the vertex stream and indices are in
  // object space; mesh m = triangles [meshFirst[m], +meshCount[m]); instance i places mesh
  // mesh[i] with the row-major 3x4 object->world affine objectToWorld[12i .. 12i+12)
  instances?: { meshFirst: Int32Array; meshCount: Int32Array; objectToWorld: Float32Array; mesh: Int32Array };
}
export declare class PathtracingRenderer {
  constructor(parameters?: { materialProfile?: 'webgl-lean' | 'webgl-full'; device?: number; width?: number; height?: number;
                             tileSchedule?: 'frames' | 'tiles' });
  maxBounces: number; useIBL: boolean; showBackground: boolean; iblRotation: number;
  backgroundColor: number[]; rayEps: number; clampThreshold: number; exposure: number;
  tonemapping: string; enableGamma: boolean; enableFxaa: boolean; pixelRatio: number; tileRes: number;
  // 'frames' (default): whole frames per native call; 'tiles': the reference's scissored tile passes
  tileSchedule: 'frames' | 'tiles';
  // debugMode other than 'None' renders the debug program (u_debug_mode = debugModes.indexOf(debugMode))
  readonly debugModes: string[]; readonly tonemappingModes: string[]; debugMode: string;
  readonly diagnostics: { backend: string; sampleCount: number; mode: string; profiling: Record<string, number> };
  setScene(scene: SceneBuffers, progress?: (e: { phase: string }) => void): Promise<void>;
  setIBL(texture: { image: { data: Float32Array; width: number; height: number } }): void;
  resize(width: number, height: number): void;
  resetAccumulation(): void;
  stopRendering(): void;
  setInteractionMode(flag: boolean): void;
  render(camera: PathtracingCamera, numSamples: number, tileFinishedCB?: () => void,
         frameFinishedCB?: (frameCount: number) => void, renderingFinishedCB?: () => void,
         framesPerCall?: number): Promise<void>;
  readAccumulation(): Float32Array;
  readDisplay(): Uint8Array;          // RGBA8 canvas of the display chain, row 0 = bottom
  saveImage(file: string): void;      // that canvas as a PNG
  counters(): Float64Array;
}
export declare function packFrameBlock(r: PathtracingRenderer, camera: PathtracingCamera, frameCount: number): Float32Array;
// renderer.ts:573-593 tile order and renderer.ts:632-643 scissor rectangle [x0, y0, width, height] (clipped; null if empty)
export declare function tileOrder(tileRes: number): number[];
export declare function tileRect(tile: number, tileRes: number, renderRes: [number, number]): [number, number, number, number] | null;

export interface RenderRequest {
  scene: number; ibl: number; camera: PathtracingCamera; width: number; height: number; spp: number;
  maxBounces?: number; useIBL?: boolean; showBackground?: boolean; backgroundColor?: number[];
  iblRotation?: number; rayEps?: number; clampThreshold?: number; debugMode?: string;
  onProgress?: (framesDone: number, framesTotal: number) => void;
}
export interface RenderResult { accumulation: Float32Array; width: number; height: number; frames: number; latencyMs: number; }
export declare class RenderServer {
  constructor(options?: { materialProfile?: 'webgl-lean' | 'webgl-full'; device?: number; memoryBudget?: number;
                          policy?: 'rr' | 'fifo' | 'srpt'; gpuBvh?: boolean });
  addScene(scene: SceneBuffers): number;
  addIbl(texture: { image: { data: Float32Array; width: number; height: number } }): number;
  submit(request: RenderRequest): Promise<RenderResult>;
  memoryUsage(): number;
}
