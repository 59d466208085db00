function pl_loadlibrary()
%PL_LOADLIBRARY  Load libparalanczos.so (MI355X) once, under the alias
%   'libparalanczos', with its plain-C header ParaLanczos.h. The library
%   lives one directory above this file, the header in include/ next to the
%   package. Replaces the reference's parallel.gpu.CUDAKernel loading of
%   PTX (Matlab/gpu_multi_subsvd.m:48-57).
if ~libisloaded('libparalanczos')
    here = fileparts(mfilename('fullpath'));
    pkg = fileparts(here);
    loadlibrary(fullfile(pkg, 'libparalanczos.so'), ...
                fullfile(fileparts(pkg), 'include', 'ParaLanczos.h'), ...
                'alias', 'libparalanczos');
end
end
