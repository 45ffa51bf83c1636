{-# LANGUAGE ForeignFunctionInterface #-}
{-# LANGUAGE BangPatterns #-}
{-|
Module      : Science.QuantumChemistry.HartreeFock.HfkFFI
Description : libhfk (MI355X ERI engine and integral-direct Fock build)
              behind the reference's Repa-facing functions.

UNCOMPILED: no GHC toolchain exists where this library is built and tested,
so this module has never been type-checked.  The C side it binds
(include/hfk.h) is exercised through ctypes by the test suite.

Drop-in use inside the reference (HartreeFock/HartreeFock.hs):

  * scfHF (:70-91): create one context per geometry with 'withHfk' and skip
    'calcIntegrals' (:82) -- the Fock build is integral-direct;
  * scfDIIS (:108): replace @fock atoms core oldDensity integrals@ by
    @fockGPU ctx core oldDensity@;
  * the exact flat array of calcIntegrals (:166-181), for small systems or
    comparisons, is 'calcIntegralsGPU'.

Every call is @safe@: calls run for milliseconds to seconds and the
@-threaded@ RTS (HartreeFock.cabal:114) must keep collecting.  Repa @U@
arrays live in unpinned memory, so inputs are copied into storable vectors
before a call and results copied back.  A failing call throws 'SCFError'
(Error.hs:19) after printing the library's message.
-}
module Science.QuantumChemistry.HartreeFock.HfkFFI
  ( HfkCtx
  , withHfk
  , calcIntegralsGPU
  , calcGmatrixGPU
  , fockGPU
  , scfGPU
  , gradientGPU
  ) where

import Control.Exception (bracket, throwIO)
import Control.Monad (when, zipWithM_)
import Data.Array.Repa (Array, DIM1, DIM2, U, Z(..), (:.)(..))
import qualified Data.Array.Repa as R
import qualified Data.Vector.Storable as VS
import qualified Data.Vector.Unboxed as VU
import Foreign
import Foreign.C.String (peekCString)
import Foreign.C.Types

import Science.QuantumChemistry.Error (HSFOCKException(..))
import Science.QuantumChemistry.GlobalTypes

-- ------------------------------------------------------------------
-- The C ABI (include/hfk.h)
-- ------------------------------------------------------------------
data HfkCtx

foreign import ccall safe "hfk_default_options" c_default_options :: Ptr () -> IO ()
foreign import ccall safe "hfk_create"     c_create     :: Ptr () -> Ptr () -> Ptr (Ptr HfkCtx) -> IO CInt
foreign import ccall safe "hfk_destroy"    c_destroy    :: Ptr HfkCtx -> IO CInt
foreign import ccall safe "hfk_eri_flat"   c_eri_flat   :: Ptr HfkCtx -> Ptr CDouble -> CSize -> IO CInt
foreign import ccall safe "hfk_fock_G"     c_fock_G     :: Ptr HfkCtx -> Ptr CDouble -> Ptr CDouble -> IO CInt
foreign import ccall safe "hfk_fock"       c_fock       :: Ptr HfkCtx -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble -> IO CInt
foreign import ccall safe "hfk_default_scf_options" c_default_scf :: Ptr () -> IO ()
foreign import ccall safe "hfk_scf"        c_scf        :: Ptr HfkCtx -> Ptr () -> Ptr () -> Ptr CDouble -> Ptr CDouble
                                                        -> Ptr CDouble -> Ptr CDouble -> IO CInt
foreign import ccall safe "hfk_gradient"   c_gradient   :: Ptr HfkCtx -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble -> IO CInt
foreign import ccall safe "hfk_last_error" c_last_error :: Ptr CChar -> CSize -> IO CInt

-- struct sizes and field offsets on x86-64 (include/hfk.h)
sizeBasis, sizeOptions, sizeScfOptions, sizeScfResult :: Int
sizeBasis = 80        -- int32 nbf, nprim_total, natom; 8 pointers from offset 16
sizeOptions = 32      -- int32 screen_mode, boys_mode; double screen_tau, schwarz_cut, fp32_threshold
sizeScfOptions = 80
sizeScfResult = 56

check :: String -> CInt -> IO ()
check what rc = when (rc /= 0) $ do
  msg <- allocaBytes 512 $ \buf -> c_last_error buf 512 >> peekCString buf
  putStrLn ("hfk: " ++ what ++ " failed (" ++ show rc ++ "): " ++ msg)
  throwIO SCFError

-- ------------------------------------------------------------------
-- [AtomData] -> struct hfk_basis: one row per CGF in [AtomData] order and
-- getBasis order, exactly the indexing of calcCoordCGF
-- (NumericalTools/LinearAlgebra.hs:293-301); Cartesian exponents from
-- mapAxesMomentum (GlobalTypes.hs:302-323); primitives (coefficient,
-- exponent) as stored, i.e. already multiplied by the primitive norm
-- (NormalizeBasis.hs:52-59).
-- ------------------------------------------------------------------
data FlatBasis = FlatBasis
  { fbCenter :: VS.Vector CDouble, fbLxyz :: VS.Vector Int32, fbOff :: VS.Vector Int32
  , fbCoef :: VS.Vector CDouble, fbExpo :: VS.Vector CDouble, fbAtom :: VS.Vector Int32
  , fbZ :: VS.Vector CDouble, fbR :: VS.Vector CDouble }

flatten :: [AtomData] -> FlatBasis
flatten atoms = FlatBasis
  { fbCenter = VS.fromList [realToFrac x | (r, _, _) <- rows, x <- r]
  , fbLxyz   = VS.fromList [fromIntegral e | (_, cgf, _) <- rows, e <- getAxesMom (getfunTyp cgf)]
  , fbOff    = VS.fromList (scanl (+) 0 [fromIntegral (length (getPrimitives cgf)) | (_, cgf, _) <- rows])
  , fbCoef   = VS.fromList [realToFrac c | (_, cgf, _) <- rows, (c, _) <- getPrimitives cgf]
  , fbExpo   = VS.fromList [realToFrac e | (_, cgf, _) <- rows, (_, e) <- getPrimitives cgf]
  , fbAtom   = VS.fromList [fromIntegral ia | (_, _, ia) <- rows]
  , fbZ      = VS.fromList [realToFrac (getZnumber a) | a <- atoms]
  , fbR      = VS.fromList [realToFrac x | a <- atoms, x <- getCoord a] }
  where rows = [(getCoord a, cgf, ia) | (ia, a) <- zip [0 :: Int ..] atoms, cgf <- getBasis a]

-- | Create a context for the molecule (reference screen at 1e-8, reference
-- Boys table, all FP64 unless fp32Threshold > 0), run the action, destroy it.
withHfk :: [AtomData] -> Double -> (Ptr HfkCtx -> IO a) -> IO a
withHfk atoms fp32Threshold act =
  let fb = flatten atoms
      nbf = VS.length (fbAtom fb)
  in VS.unsafeWith (fbCenter fb) $ \pc -> VS.unsafeWith (fbLxyz fb) $ \pl ->
     VS.unsafeWith (fbOff fb) $ \po -> VS.unsafeWith (fbCoef fb) $ \pcf ->
     VS.unsafeWith (fbExpo fb) $ \pe -> VS.unsafeWith (fbAtom fb) $ \pa ->
     VS.unsafeWith (fbZ fb) $ \pz -> VS.unsafeWith (fbR fb) $ \pr ->
     allocaBytes sizeBasis $ \b -> allocaBytes sizeOptions $ \o -> alloca $ \pctx -> do
       pokeByteOff b 0 (fromIntegral nbf :: Int32)
       pokeByteOff b 4 (fromIntegral (VS.length (fbCoef fb)) :: Int32)
       pokeByteOff b 8 (fromIntegral (length atoms) :: Int32)
       zipWithM_ (pokeByteOff b) [16, 24 ..] [castPtr pc, castPtr pl, castPtr po, castPtr pcf, castPtr pe,
                                              castPtr pa, castPtr pz, castPtr pr :: Ptr ()]
       c_default_options o
       pokeByteOff o 24 (realToFrac fp32Threshold :: CDouble)
       bracket (c_create b o pctx >>= check "hfk_create" >> peek pctx)
               (\ctx -> c_destroy ctx)
               act

-- ------------------------------------------------------------------
-- The reference's functions
-- ------------------------------------------------------------------
toStorable :: Array U DIM1 Double -> VS.Vector CDouble
toStorable = VS.map realToFrac . VS.convert . R.toUnboxed

fromStorable :: VS.Vector CDouble -> Array U DIM1 Double
fromStorable v = R.fromUnboxed (Z :. VS.length v) (VU.convert (VS.map realToFrac v))

nbfOf :: Int -> Int   -- n from the flat length n(n+1)/2
nbfOf m = floor ((sqrt (8 * fromIntegral m + 1 :: Double) - 1) / 2)

-- | calcIntegrals (HartreeFock.hs:166-181): every canonical ERI in
-- fourIndex2Flat order, M(M+1)/2 doubles (M = n(n+1)/2).
calcIntegralsGPU :: Ptr HfkCtx -> Int -> IO (Array U DIM1 Double)
calcIntegralsGPU ctx n = do
  let m = n * (n + 1) `div` 2
      len = m * (m + 1) `div` 2
  fp <- mallocForeignPtrArray len
  withForeignPtr fp $ \p -> c_eri_flat ctx p (fromIntegral len) >>= check "hfk_eri_flat"
  return (fromStorable (VS.unsafeFromForeignPtr0 fp len))

-- | calcGmatrix (HartreeFock.hs:225-241), integral-direct: the Integrals
-- argument of the reference is not needed.
calcGmatrixGPU :: Ptr HfkCtx -> Array U DIM1 Double -> IO (Array U DIM1 Double)
calcGmatrixGPU ctx density = do
  let dv = toStorable density
      m = VS.length dv
  fp <- mallocForeignPtrArray m
  VS.unsafeWith dv $ \pd -> withForeignPtr fp $ \pg -> c_fock_G ctx pd pg >>= check "hfk_fock_G"
  return (fromStorable (VS.unsafeFromForeignPtr0 fp m))

-- | fock (HartreeFock.hs:213-222): F = H + G(D), flat in and out.
fockGPU :: Ptr HfkCtx -> Array U DIM1 Double -> Array U DIM1 Double -> IO (Array U DIM1 Double)
fockGPU ctx core density = do
  let hv = toStorable core
      dv = toStorable density
      m = VS.length dv
  fp <- mallocForeignPtrArray m
  VS.unsafeWith hv $ \ph -> VS.unsafeWith dv $ \pd -> withForeignPtr fp $ \pf ->
    c_fock ctx ph pd pf >>= check "hfk_fock"
  return (fromStorable (VS.unsafeFromForeignPtr0 fp m))

-- | scfHF (HartreeFock.hs:70-123) with the reference's defaults (core
-- guess, 30 steps, 5 DIIS vectors switched on at max|e| < 0.1, converged at
-- max|e| < 1e-7), all on the GPU: HFData of GlobalTypes.hs:122-128.
scfGPU :: Ptr HfkCtx -> Int -> Double -> IO HFData
scfGPU ctx n charge =
  allocaBytes sizeScfOptions $ \o -> allocaBytes sizeScfResult $ \res -> do
    c_default_scf o
    pokeByteOff o 32 (realToFrac charge :: CDouble)   -- hfk_scf_options.charge
    let m = n * (n + 1) `div` 2
    fc <- mallocForeignPtrArray (n * n)
    fe <- mallocForeignPtrArray n
    fd <- mallocForeignPtrArray m
    ff <- mallocForeignPtrArray m
    withForeignPtr fc $ \pc -> withForeignPtr fe $ \pe -> withForeignPtr fd $ \pd -> withForeignPtr ff $ \pf ->
      c_scf ctx o res pc pe pd pf >>= check "hfk_scf"
    energy <- peekByteOff res 0 :: IO CDouble
    let coeff = R.fromUnboxed (Z :. n :. n) (VU.convert (VS.map realToFrac (VS.unsafeFromForeignPtr0 fc (n * n))))
                  :: Array U DIM2 Double
        orbE = VU.convert (VS.map realToFrac (VS.unsafeFromForeignPtr0 fe n)) :: VU.Vector Double
    return $ HFData (fromStorable (VS.unsafeFromForeignPtr0 ff m)) coeff (fromStorable (VS.unsafeFromForeignPtr0 fd m))
                    orbE (realToFrac energy)

-- | Nuclear gradient dE/dR at a converged density (no reference
-- counterpart: HartreeFock/Derivatives.hs does not compile); 3 natom
-- values, atom-major, Eh/bohr.
gradientGPU :: Ptr HfkCtx -> Int -> Array U DIM1 Double -> IO (Array U DIM1 Double)
gradientGPU ctx natom density = do
  let dv = toStorable density
  fp <- mallocForeignPtrArray (3 * natom)
  VS.unsafeWith dv $ \pd -> withForeignPtr fp $ \pg -> c_gradient ctx pd nullPtr pg >>= check "hfk_gradient"
  return (fromStorable (VS.unsafeFromForeignPtr0 fp (3 * natom)))
