/*
 * digHolo-compatible C API of the MI355X batch-processing library
 * (libdigholo.so, built from digholo_amd/csrc).
 *
 * Every declaration below keeps the signature and argument meaning of the
 * reference declaration it replaces; the citation after each one is
 * reference digHolo.h:line.  Unlike the reference on Linux (whose EXT_C is
 * empty, digHolo.h:148-156, so its .so exports C++-mangled names), these are
 * declared with C linkage; libdigholo.so exports every function under both
 * names, so a C++ program built against the reference header links too.
 * Defining DIGHOLO_CXX_LINKAGE before this header declares them with C++
 * linkage, as the reference header does.
 *
 * Extensions with no reference counterpart are grouped at the end and say so.
 */
#ifndef DIGHOLO_MI355X_H
#define DIGHOLO_MI355X_H

#if defined(__cplusplus) && !defined(DIGHOLO_CXX_LINKAGE)
#define EXT_C extern "C"
#else
#define EXT_C
#endif

typedef float complex64[2];                       /* digHolo.h:163 */

/* Error codes: digHolo.h:171-182 */
#define DIGHOLO_ERROR_SUCCESS 0
#define DIGHOLO_ERROR_ERROR 1
#define DIGHOLO_ERROR_INVALIDHANDLE 2
#define DIGHOLO_ERROR_NULLPOINTER 3
#define DIGHOLO_ERROR_SETFRAMEBUFFERDISABLED 4
#define DIGHOLO_ERROR_INVALIDDIMENSION 5
#define DIGHOLO_ERROR_INVALIDPOLARISATION 6
#define DIGHOLO_ERROR_INVALIDAXIS 7
#define DIGHOLO_ERROR_INVALIDARGUMENT 8
#define DIGHOLO_ERROR_MEMORYALLOCATION 9
#define DIGHOLO_ERROR_FILENOTCREATED 10
#define DIGHOLO_ERROR_FILENOTFOUND 11

/* Averaging modes: digHolo.h:655-658 */
#define DIGHOLO_AVGMODE_SEQUENTIAL 0
#define DIGHOLO_AVGMODE_INTERLACED 1
#define DIGHOLO_AVGMODE_SEQUENTIALSWEEP 2
#define DIGHOLO_AVGMODE_COUNT 3

/* Wavelength ordering: digHolo.h:874-878; wavelength unit: digHolo.h:191 */
#define DIGHOLO_WAVELENGTHORDER_INPUT 0
#define DIGHOLO_WAVELENGTHORDER_OUTPUT 1
#define DIGHOLO_WAVELENGTHORDER_FAST 0
#define DIGHOLO_WAVELENGTHORDER_SLOW 1
#define DIGHOLO_UNIT_LAMBDA 1

/* AutoAlign modes: digHolo.h:925-928 */
#define DIGHOLO_AUTOALIGNMODE_FULL 0
#define DIGHOLO_AUTOALIGNMODE_TWEAK 1
#define DIGHOLO_AUTOALIGNMODE_ESTIMATE 2
#define DIGHOLO_AUTOALIGNMODE_COUNT 3

/* Basis types: digHolo.h:1348-1350 */
#define DIGHOLO_BASISTYPE_HG 0
#define DIGHOLO_BASISTYPE_LG 1
#define DIGHOLO_BASISTYPE_CUSTOM 2

/* Power-meter metrics: digHolo.h:1363-1372 */
#define DIGHOLO_METRIC_IL 0
#define DIGHOLO_METRIC_MDL 1
#define DIGHOLO_METRIC_DIAG 2
#define DIGHOLO_METRIC_SNRAVG 3
#define DIGHOLO_METRIC_DIAGBEST 4
#define DIGHOLO_METRIC_DIAGWORST 5
#define DIGHOLO_METRIC_SNRBEST 6
#define DIGHOLO_METRIC_SNRWORST 7
#define DIGHOLO_METRIC_SNRMG 8
#define DIGHOLO_METRIC_COUNT 9

/* Analysis parameters: digHolo.h:1624-1631 (MAXABSIDX holds int32 bits) */
#define DIGHOLO_ANALYSIS_TOTALPOWER 0
#define DIGHOLO_ANALYSIS_COMX 1
#define DIGHOLO_ANALYSIS_COMY 2
#define DIGHOLO_ANALYSIS_MAXABS 3
#define DIGHOLO_ANALYSIS_MAXABSIDX 4
#define DIGHOLO_ANALYSIS_AEFF 5
#define DIGHOLO_ANALYSIS_COMYWRAP 6
#define DIGHOLO_ANALYSIS_COUNT 7

/* Benchmark info indices: digHolo.h:1783-1788 */
#define DIGHOLO_BENCHMARK_FFT 0
#define DIGHOLO_BENCHMARK_IFFT 1
#define DIGHOLO_BENCHMARK_APPLYTILT 2
#define DIGHOLO_BENCHMARK_BASIS 3
#define DIGHOLO_BENCHMARK_OVERLAP 4
#define DIGHOLO_BENCHMARK_TOTAL 5

/* Verbosity: digHolo.h:1906-1909 */
#define DIGHOLO_VERBOSITY_OFF 0
#define DIGHOLO_VERBOSITY_BASIC 1
#define DIGHOLO_VERBOSITY_DEBUG 2
#define DIGHOLO_VERBOSITY_COOKED 3

/* ---- lifecycle ---- */
EXT_C int digHoloCreate(void);                                                    /* digHolo.h:209 */
EXT_C int digHoloDestroy(int handleIdx);                                          /* digHolo.h:216 */

/* ---- frame buffer ---- */
EXT_C int digHoloSetFrameBuffer(int handleIdx, float* buffer);                    /* digHolo.h:245 */
EXT_C float* digHoloGetFrameBuffer(int handleIdx);                                /* digHolo.h:255 */
EXT_C int digHoloSetFrameBufferUint16(int handleIdx, unsigned short* buffer, int transposeMode); /* digHolo.h:268 */
EXT_C unsigned short* digHoloGetFrameBufferUint16(int handleIdx, int* transposeMode);            /* digHolo.h:277 */

/* ---- camera / frame configuration ---- */
EXT_C int digHoloConfigSetFrameDimensions(int handleIdx, int width, int height);  /* digHolo.h:298 */
EXT_C int digHoloConfigGetFrameDimensions(int handleIdx, int* width, int* height);/* digHolo.h:308 */
EXT_C int digHoloConfigSetFramePixelSize(int handleIdx, float pixelSize);         /* digHolo.h:317 */
EXT_C float digHoloConfigGetFramePixelSize(int handleIdx);                        /* digHolo.h:325 */
EXT_C int digHoloConfigSetPolCount(int handleIdx, int polCount);                  /* digHolo.h:334 */
EXT_C int digHoloConfigGetPolCount(int handleIdx);                                /* digHolo.h:344 */
EXT_C int digHoloConfigSetFillFactorCorrectionEnabled(int handleIdx, int enabled);/* digHolo.h:449 */
EXT_C int digHoloConfigGetFillFactorCorrectionEnabled(int handleIdx);             /* digHolo.h:460 */

/* ---- batch ---- */
EXT_C int digHoloSetBatch(int handleIdx, int batchCount, float* frameBuffer);     /* digHolo.h:483 */
EXT_C int digHoloSetBatchAvg(int handleIdx, int batchCount, float* frameBuffer, int avgCount, int avgMode); /* digHolo.h:502 */
EXT_C int digHoloSetBatchUint16(int handleIdx, int batchCount, unsigned short* frameBuffer, int transpose); /* digHolo.h:519 */
EXT_C int digHoloSetBatchAvgUint16(int handleIdx, int batchCount, unsigned short* frameBuffer, int avgCount, int avgMode, int transpose); /* digHolo.h:540 */
EXT_C int digHoloConfigSetBatchCount(int handleIdx, int batchCount);              /* digHolo.h:551 */
EXT_C int digHoloConfigSetBatchCalibration(int handleIdx, complex64* cal, int polCount, int batchCount); /* digHolo.h:599 */
EXT_C complex64* digHoloConfigGetBatchCalibration(int handleIdx, int* polCount, int* batchCount);       /* digHolo.h:611 */
EXT_C int digHoloConfigSetBatchCalibrationFromFile(int handleIdx, const char* fname, int polCount, int batchCount); /* digHolo.h:624 */
EXT_C int digHoloConfigSetBatchCalibrationEnabled(int handleIdx, int enabled);    /* digHolo.h:635 */
EXT_C int digHoloConfigGetBatchCalibrationEnabled(int handleIdx);                 /* digHolo.h:645 */
/* Reference-wave calibration: frames (wavelengthCount x height x width) of the
   reference wave alone, as intensity or field; applied, low-passed through the
   field's own Fourier window, to every reconstructed field whose frame has the
   same dimensions.  A calibration of one frame serves every wavelength. */
EXT_C int digHoloConfigSetRefCalibrationIntensity(int handleIdx, unsigned short* cal, int wavelengthCount, int width,
                                                  int height);                      /* digHolo.h:369 */
EXT_C int digHoloConfigSetRefCalibrationField(int handleIdx, complex64* cal, int wavelengthCount, int width,
                                              int height);                          /* digHolo.h:384 */
EXT_C int digHoloConfigSetRefCalibrationFromFile(int handleIdx, const char* fname, int wavelengthCount, int width,
                                                 int height);                       /* digHolo.h:400 */
EXT_C int digHoloConfigSetRefCalibrationEnabled(int handleIdx, int enabled);      /* digHolo.h:409 */
EXT_C int digHoloConfigGetRefCalibrationEnabled(int handleIdx);                   /* digHolo.h:417 */
EXT_C complex64* digHoloConfigGetRefCalibrationFields(int handleIdx, int* wavelengthCount, int* polCount, float** x,
                                                      float** y, int* width, int* height); /* digHolo.h:437 */
EXT_C int digHoloConfigGetBatchCount(int handleIdx);                              /* digHolo.h:559 */
EXT_C int digHoloConfigSetBatchAvgCount(int handleIdx, int avgCount);             /* digHolo.h:570 */
EXT_C int digHoloConfigGetBatchAvgCount(int handleIdx);                           /* digHolo.h:580 */
EXT_C int digHoloConfigSetBatchAvgMode(int handleIdx, int avgMode);               /* digHolo.h:670 */
EXT_C int digHoloConfigGetBatchAvgMode(int handleIdx);                            /* digHolo.h:680 */

/* ---- FFT / IFFT windows ---- */
EXT_C int digHoloConfigSetfftWindowSize(int handleIdx, int width, int height);    /* digHolo.h:701 */
EXT_C int digHoloConfigGetfftWindowSize(int handleIdx, int* width, int* height);  /* digHolo.h:712 */
EXT_C int digHoloConfigSetfftWindowSizeX(int handleIdx, int width);               /* digHolo.h:722 */
EXT_C int digHoloConfigGetfftWindowSizeX(int handleIdx);                          /* digHolo.h:730 */
EXT_C int digHoloConfigGetfftWindowSizeY(int handleIdx);                          /* digHolo.h:739 */
EXT_C int digHoloConfigSetfftWindowSizeY(int handleIdx, int height);              /* digHolo.h:749 */
EXT_C int digHoloConfigSetFourierWindowRadius(int handleIdx, float windowRadius); /* digHolo.h:766 */
EXT_C float digHoloConfigGetFourierWindowRadius(int handleIdx);                   /* digHolo.h:774 */
EXT_C int digHoloConfigSetIFFTResolutionMode(int handleIdx, int resolutionModeIdx);/* digHolo.h:786 */
EXT_C int digHoloConfigGetIFFTResolutionMode(int handleIdx);                      /* digHolo.h:797 */
EXT_C int digHoloConfigSetWavelengthCentre(int handleIdx, float lambda0);         /* digHolo.h:814 */
EXT_C float digHoloConfigGetWavelengthCentre(int handleIdx);                      /* digHolo.h:822 */
/* Wavelength sweeps: frames at L wavelengths, wavelength the fast (0) or slow
   (1) axis of the input frames and of the output fields/coefficients.  A slow
   ordering requires batchCount to be a multiple of L; frame averaging is not
   available together with a sweep of more than one wavelength. */
EXT_C int digHoloConfigSetWavelengths(int handleIdx, float* wavelengths, int wavelengthCount); /* digHolo.h:832 */
EXT_C int digHoloConfigSetWavelengthsLinearFrequency(int handleIdx, float wavelengthStart, float wavelengthStop,
                                                      int wavelengthCount);          /* digHolo.h:843 */
EXT_C float* digHoloConfigGetWavelengths(int handleIdx, int* wavelengthCount);     /* digHolo.h:854 */
EXT_C int digHoloConfigSetWavelengthOrdering(int handleIdx, int inoutIdx, int ordering); /* digHolo.h:886 */
EXT_C int digHoloConfigGetWavelengthOrdering(int handleIdx, int inoutIdx);         /* digHolo.h:895 */

/* ---- AutoAlign switches ---- */
EXT_C int digHoloConfigSetAutoAlignBeamCentre(int handleIdx, int enable);         /* digHolo.h:938 */
EXT_C int digHoloConfigGetAutoAlignBeamCentre(int handleIdx);                     /* digHolo.h:946 */
EXT_C int digHoloConfigSetAutoAlignDefocus(int handleIdx, int enable);            /* digHolo.h:955 */
EXT_C int digHoloConfigGetAutoAlignDefocus(int handleIdx);                        /* digHolo.h:963 */
EXT_C int digHoloConfigSetAutoAlignTilt(int handleIdx, int enable);               /* digHolo.h:972 */
EXT_C int digHoloConfigGetAutoAlignTilt(int handleIdx);                           /* digHolo.h:980 */
EXT_C int digHoloConfigSetAutoAlignBasisWaist(int handleIdx, int enable);         /* digHolo.h:989 */
EXT_C int digHoloConfigGetAutoAlignBasisWaist(int handleIdx);                     /* digHolo.h:997 */
EXT_C int digHoloConfigSetAutoAlignFourierWindowRadius(int handleIdx, int enable);/* digHolo.h:1008 */
EXT_C int digHoloConfigGetAutoAlignFourierWindowRadius(int handleIdx);            /* digHolo.h:1016 */
EXT_C int digHoloConfigSetAutoAlignTol(int handleIdx, float tolerance);           /* digHolo.h:1025 */
EXT_C float digHoloConfigGetAutoAlignTol(int handleIdx);                          /* digHolo.h:1033 */
EXT_C int digHoloConfigSetAutoAlignPolIndependence(int handleIdx, int polIndependence); /* digHolo.h:1046 */
EXT_C int digHoloConfigGetAutoAlignPolIndependence(int handleIdx);                /* digHolo.h:1058 */
EXT_C int digHoloConfigSetAutoAlignBasisMulConjTrans(int handleIdx, int mulConjTranspose); /* digHolo.h:1071 */
EXT_C int digHoloConfigGetAutoAlignBasisMulConjTrans(int handleIdx);              /* digHolo.h:1083 */
EXT_C int digHoloConfigSetAutoAlignMode(int handleIdx, int alignMode);            /* digHolo.h:1101 */
EXT_C int digHoloConfigGetAutoAlignMode(int handleIdx);                           /* digHolo.h:1109 */
EXT_C int digHoloConfigSetAutoAlignGoalIdx(int handleIdx, int goalIdx);           /* digHolo.h:1128 */
EXT_C int digHoloConfigGetAutoAlignGoalIdx(int handleIdx);                        /* digHolo.h:1136 */

/* ---- AutoAlign ---- */
/* FULL (estimate from the frames, then tweak), TWEAK (search from the current
   values) and ESTIMATE (estimate only, then one evaluation at the whole-half-
   frame FFT window it used); returns the goal metric in dB (-FLT_MAX on error,
   or with no basis, where the power meter has no modes to measure). */
EXT_C float digHoloAutoAlign(int handleIdx);                                      /* digHolo.h:1152 */
EXT_C float digHoloAutoAlignGetMetric(int handleIdx, int metricIdx);              /* digHolo.h:1163 */
EXT_C float* digHoloAutoAlignGetMetrics(int handleIdx, int metricIdx);            /* digHolo.h:1174 */
EXT_C int digHoloAutoAlignCalcMetrics(int handleIdx);                             /* digHolo.h:1185 */

/* ---- alignment parameters ---- */
EXT_C int digHoloConfigSetBeamCentre(int handleIdx, int axisIdx, int polIdx, float centre); /* digHolo.h:1205 */
EXT_C float digHoloConfigGetBeamCentre(int handleIdx, int axisIdx, int polIdx);   /* digHolo.h:1215 */
EXT_C int digHoloConfigSetTilt(int handleIdx, int axisIdx, int polIdx, float tilt);/* digHolo.h:1226 */
EXT_C float digHoloConfigGetTilt(int handleIdx, int axisIdx, int polIdx);         /* digHolo.h:1236 */
EXT_C int digHoloConfigSetDefocus(int handleIdx, int polIdx, float defocus);      /* digHolo.h:1246 */
EXT_C float digHoloConfigGetDefocus(int handleIdx, int polIdx);                   /* digHolo.h:1255 */
EXT_C int digHoloConfigSetPolLockTilt(int handleIdx, int polLock);                /* digHolo.h:1270 */
EXT_C int digHoloConfigGetPolLockTilt(int handleIdx);                             /* digHolo.h:1282 */
EXT_C int digHoloConfigSetPolLockDefocus(int handleIdx, int polLock);             /* digHolo.h:1297 */
EXT_C int digHoloConfigGetPolLockDefocus(int handleIdx);                          /* digHolo.h:1309 */
EXT_C int digHoloConfigSetPolLockBasisWaist(int handleIdx, int polLock);          /* digHolo.h:1318 */
EXT_C int digHoloConfigGetPolLockBasisWaist(int handleIdx);                       /* digHolo.h:1326 */

/* ---- basis ---- */
EXT_C int digHoloConfigSetBasisGroupCount(int handleIdx, int groupCount);         /* digHolo.h:1384 */
EXT_C int digHoloConfigGetBasisGroupCount(int handleIdx);                         /* digHolo.h:1393 */
EXT_C int digHoloConfigSetBasisWaist(int handleIdx, int polIdx, float waist);     /* digHolo.h:1404 */
EXT_C float digHoloConfigGetBasisWaist(int handleIdx, int polIdx);                /* digHolo.h:1416 */
EXT_C int digHoloConfigSetBasisTypeHG(int handleIdx);                             /* digHolo.h:1424 */
EXT_C int digHoloConfigSetBasisTypeLG(int handleIdx);                             /* digHolo.h:1435 */
EXT_C int digHoloConfigSetBasisTypeCustom(int handleIdx, int modeCountIn, int modeCountOut,
                                          complex64* transform);                 /* digHolo.h:1453 */
EXT_C int digHoloConfigSetBasisType(int handleIdx, int basisType);                /* digHolo.h:1464 */
EXT_C int digHoloConfigGetBasisType(int handleIdx);                               /* digHolo.h:1472 */

/* ---- processing ---- */
EXT_C complex64* digHoloProcessBatch(int handleIdx, int* batchCount, int* modeCount, int* polCount); /* digHolo.h:1494 */
EXT_C complex64* digHoloProcessBatchFrequencySweepLinear(int handleIdx, int* batchCount, int* modeCount, int* polCount,
                                                         float lambdaStart, float lambdaStop,
                                                         int lambdaCount);           /* digHolo.h:1510 */
EXT_C complex64* digHoloProcessBatchWavelengthSweepArbitrary(int handleIdx, int* batchCount, int* modeCount,
                                                             int* polCount, float* wavelengths,
                                                             int lambdaCount);       /* digHolo.h:1525 */
EXT_C complex64* digHoloGetFields(int handleIdx, int* batchCount, int* polCount, float** x, float** y, int* width, int* height); /* digHolo.h:1549 */
EXT_C complex64* digHoloBasisGetCoefs(int handleIdx, int* batchCount, int* modeCount, int* polCount); /* digHolo.h:1603 */
EXT_C complex64* digHoloBasisGetFields(int handleIdx, int* modeCount, int* polCount, float** x, float** y,
                                       int* width, int* height);                 /* digHolo.h:1587 */
/* int16 fields R, I [batch][pol][h][w] with per (pol, batch) scale 32767 / max|V| */
EXT_C int digHoloGetFields16(int handleIdx, int* batchCount, int* polCount, short** fieldR, short** fieldI,
                             float** fieldScale, float** x, float** y, int* width, int* height); /* digHolo.h:1568 */
/* planeIdx 0: R2C Fourier plane of the frames, 1: reconstructed fields.  parameters
   [DIGHOLO_ANALYSIS_COUNT][pol][batch + 1] (last: the batch-summed intensity),
   totalIntensity [pol][W * H] (pixelCountY x pixelCountX used per pol).  Computed on
   request from the last batch's device frames and fields. */
EXT_C int digHoloBatchGetSummary(int handleIdx, int planeIdx, int* parameterCount, int* totalCount, int* polCount,
                                 float** parameters, int* pixelCountX, int* pixelCountY, float** xAxis,
                                 float** yAxis, float** totalIntensity);           /* digHolo.h:1659 */

/* ---- threads / console ---- */
/* The processing pipeline one step at a time (digHoloProcessBatch runs all
 * four).  FFT: frames -> Fourier plane of the window columns; IFFT: Fourier
 * window -> field still carrying the reference tilt; RemoveTilt: the final
 * field (GetFields valid); BasisExtractCoefs: coefficients of the current
 * fields (BasisGetCoefs valid).  A step whose input is missing returns
 * DIGHOLO_ERROR_NULLPOINTER. */
EXT_C int digHoloProcessFFT(int handleIdx);                                     /* digHolo.h:1670 */
EXT_C int digHoloProcessIFFT(int handleIdx);                                    /* digHolo.h:1682 */
EXT_C int digHoloProcessRemoveTilt(int handleIdx);                              /* digHolo.h:1694 */
EXT_C int digHoloProcessBasisExtractCoefs(int handleIdx);                       /* digHolo.h:1706 */

EXT_C int digHoloConfigSetThreadCount(int handleIdx, int threadCount);            /* digHolo.h:1749 */
EXT_C int digHoloConfigGetThreadCount(int handleIdx);                             /* digHolo.h:1757 */
/* Returns batches per second end to end (digHolo.cpp:13345, :13396); info
 * must hold 7 floats: per-stage calls per second at the indices below, [5] the
 * rate the stage times add up to, [6] the measured end-to-end rate
 * (digHolo.cpp:13374-13382). */
EXT_C float digHoloBenchmark(int handleIdx, float goalDuration, float* info);     /* digHolo.h:1776 */
EXT_C int digHoloConfigSetVerbosity(int handleIdx, int verbosity);                /* digHolo.h:1924 */
EXT_C int digHoloConfigGetVerbosity(int handleIdx);                               /* digHolo.h:1937 */

/* =====================================================================
 * Extensions (no counterpart in the reference)
 * ===================================================================== */

/* Overlap precision: 0 = plain FP32, 1 = split-FP16 MFMA (hi+lo operands,
 * FP32 accumulate).  Default 0. */
#define DIGHOLO_PRECISION_FP32 0
#define DIGHOLO_PRECISION_SPLIT16 1
EXT_C int digHoloConfigSetOverlapPrecision(int handleIdx, int precision);
EXT_C int digHoloConfigGetOverlapPrecision(int handleIdx);

/* Same work as digHoloProcessBatch, but asynchronous on the handle's stream and
 * returning a DEVICE pointer to the coefficients [batch][pol][mode]; the
 * caller synchronises (digHoloSynchronize) before reading it.  Frames set
 * with SetBatch* may be host or device pointers. */
EXT_C complex64* digHoloProcessBatchDevice(int handleIdx, int* batchCount, int* modeCount, int* polCount);
/* Device pointer to the reconstructed fields [batch][pol][height][width]. */
EXT_C complex64* digHoloGetFieldsDevice(int handleIdx, int* batchCount, int* polCount, int* width, int* height);
/* Blocks until all work queued on the handle's stream is done. */
EXT_C int digHoloSynchronize(int handleIdx);
/* Runs the work of the handle on an existing HIP stream (hipStream_t passed as
 * void*); NULL reverts to the handle's own stream. */
EXT_C int digHoloSetStream(int handleIdx, void* stream);
EXT_C void* digHoloGetStream(int handleIdx);

/* Window geometry derived from the current configuration, without touching
 * the GPU: out[0..11] = W, H, w, h, cx0, cy0, cx1, cy1, SX0, SY0, SX1, SY1. */
EXT_C int digHoloGetWindowGeometry(int handleIdx, int* out);

/* Window path.  mode 3 (default): the size-specialised camera-window-to-field
 * kernel (row transforms on MFMA for 256-wide FFT windows, in registers for
 * the others; column transform as a split-fp16 GEMM on MFMA; uint16 frames,
 * W/2 in {64, 96, 128}); 2: the general fused kernel with
 * the MFMA column transform (uint16 frames); 1: the fused kernel with FFT column
 * transforms; 0: the staged kernels.  A mode the shape or input type does not
 * fit falls back to the next lower one.  GetWindowPath reports the path of the
 * last batch (3, 2, 1, 0; -1 none yet). */
EXT_C int digHoloConfigSetFusedWindow(int handleIdx, int enabled);
/* Overlap inside the specialised window kernel: -1 (default) for the split-fp16
 * form only, 0 never (the separate overlap kernel), 1 in either precision. */
EXT_C int digHoloConfigSetFusedOverlap(int handleIdx, int mode);
EXT_C int digHoloGetWindowPath(int handleIdx);
/* Row transform of the size-specialised window kernel for 256-wide FFT windows:
 * 0 (default) on MFMA (a 16-point DFT of the samples, then one matrix per
 * residue pair for the rest of the transform and the r2c unpack), 1 in
 * registers (two FFT passes).  Other widths always use the register
 * transform.  The autotuner times both.  No counterpart in the reference. */
EXT_C int digHoloConfigSetWindowRowTransform(int handleIdx, int mode);
EXT_C int digHoloConfigGetWindowRowTransform(int handleIdx);

/* Per-stage GPU time in ms, measured with HIP events on the handle's stream
 * when enabled, averaged over every batch processed since the previous call:
 * ms[0..4] = row FFT (the whole fused window kernel on the fused path),
 * column FFT + Fourier window, IFFT + tilt (both 0 on the fused path),
 * HG basis/tables (0 when cached), overlap. */
EXT_C int digHoloConfigSetStageTiming(int handleIdx, int enabled);
EXT_C int digHoloGetStageTimes(int handleIdx, float* ms);

/* Autotuning of the launch variants per batch shape (frame and window sizes,
 * pols, groups, precision, input type, batch size rounded up to a power of
 * two): window path, split overlap fused or separate, overlap tiles per wave,
 * each timed on the batch with HIP events.  mode 0 (default): default kernels;
 * 1: tune each new shape on first use and remember the fastest (in memory and
 * in the cache file $DIGHOLO_TUNE_CACHE, else $HOME/.cache/digholo_amd_tune.txt);
 * 2: use remembered choices only. */
EXT_C int digHoloConfigSetAutoTune(int handleIdx, int mode);
EXT_C int digHoloConfigGetAutoTune(int handleIdx);

/* Processes the current batch of each listed handle (distinct handles, device-
 * resident frames) on the first handle's stream.  The first call for a handle
 * list captures the launches into a hipGraph: batches on the size-specialised
 * window kernel (HG basis, square windows) share one launch of the multi-batch
 * window kernel when they hold at most 1024 (pol, frame) items together, else
 * run one after the other; other batches are branches of the graph.  Later
 * calls repeat the launches (directly when the graph holds nothing else, else
 * by replaying it) while every handle's configuration, frames and kernels are
 * unchanged; otherwise they re-capture.  Work queued through the API on another
 * listed handle's stream before the call runs first, and that handle's next
 * API call follows it.  Outputs as digHoloProcessBatchDevice /
 * digHoloGetFieldsDevice of each handle; synchronise the first handle's stream
 * (digHoloSynchronize) before reading them. */
EXT_C int digHoloProcessBatchMultiDevice(int count, const int* handleIdx);

/* Allocation-free batch (no counterpart in the reference), for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops.  Processes batchCount device frames (uint16 when frameIsUint16, else
 * float32; frameWidth x frameHeight each, 8-byte aligned) with the
 * configuration whose tables the handle built last (one earlier
 * digHoloProcessBatch* call with it is required; settings changed since are not
 * seen) into the caller's device buffers: fields [B][pol][h][w] and, with a
 * basis, HG coefficients [B][pol][modes] (the configured overlap precision).
 * Asynchronous on the handle's stream.  It calls no allocator: the staged
 * kernels keep their intermediates in a fixed 64 MiB device array per GPU
 * (B x pol x (H + h) x w x 8 bytes; a larger batch returns
 * DIGHOLO_ERROR_MEMORYALLOCATION and is not processed).  Every call on one GPU
 * shares that array: calls that overlap on one GPU, from any handle or stream,
 * race.  One wavelength, IFFTResolutionMode 1, HG basis, no frame averaging or
 * transposition. */
EXT_C int digHoloProcessBatchDeviceNoAlloc(int handleIdx, const void* frames, int batchCount, int frameIsUint16,
                                           complex64* fields, complex64* coefs);

/* Alignment-loss gradient.  Processes the batch at the current configuration
 * and returns loss = -(1/B) sum |c|^2 + mdlWeight (ln s_max^2 - ln s_min^2)
 * (minus the linear insertion loss without the SVD, plus a mode-dependent-loss
 * term over the singular values s of the coefficient matrix the MDL metric
 * uses) and grad[pol * 6 + k] = d loss / d parameter k of pol, k in the
 * AutoAlign order waist (m), defocus (dioptre), tilt x (degree), centre x (m),
 * tilt y (degree), centre y (m), from the adjoint of the path (exact for the
 * current window placement; pol-locked parameters collect both pols' terms on
 * pol 0).  holdPlacement != 0 keeps the FFT crops and Fourier window (position,
 * size, mask) of the previous call, so the loss is a smooth function of the
 * parameters between calls (for finite differences and line searches). */
EXT_C int digHoloAutoAlignGradient(int handleIdx, int holdPlacement, float mdlWeight, float* loss, float* grad);
/* Gradient-based alignment beside digHoloAutoAlign's search: BFGS on the
 * loss above (mdlWeight as there) over the parameters the AutoAlign enable
 * flags select (pol locks respected), from the current values, at most
 * maxIterations iterations (<= 0: 30).  Ends processed
 * at the best point with its metrics stored (AutoAlignGetMetric); returns the
 * goal metric in dB. */
EXT_C float digHoloAutoAlignGradientSearch(int handleIdx, int maxIterations, float mdlWeight);

#endif
